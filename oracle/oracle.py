"""ctypes binding of the CPU oracle (oracle/libctoracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / reported CPU baseline.  The
product library never imports this module.

All functions take and return numpy arrays on the host.  ``prec`` selects the
FP32 (32) or FP64 (64) instantiation of the restatement.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

from gpu_amd.geometry import CtGeometry, Geometry

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libctoracle.so"


def build() -> Path:
    r = subprocess.run(["make", "-s", "-B", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def _load() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(LIB_PATH))
    P, G, I, D, I64 = ctypes.c_void_p, ctypes.POINTER(CtGeometry), ctypes.c_int, ctypes.c_double, ctypes.c_int64
    sig = {
        "orc_num_threads": ([], I),
        "orc_set_threads": ([I], None),
        "orc_view_angle": ([G, I, I, I, ctypes.POINTER(D), ctypes.POINTER(D)], None),
        "orc_forward": ([G, I, P, P, I, I], None),
        "orc_sart_correction": ([G, I, P, P, P, I, I], None),
        "orc_backproject_voxel": ([G, I, P, P, I, I, D, I], None),
        "orc_sart_sweep": ([G, I, P, P, D, I, I], None),
        "orc_adjoint": ([G, I, P, P, I, I], None),
        "orc_ray_segments": ([G, I, I, I, I, I, P, P, I], I),
        "orc_fista_momentum": ([I, P, P, P, I64, D], None),
        "orc_tv_fgp": ([I, I, I, I, P, P, D, I], None),
        "orc_tv_fgp_duals": ([I, I, I, I, P, P, P, D, I], None),
        "orc_tv_D": ([I, I, I, I, P, P, P, P], None),
        "orc_tv_Dt": ([I, I, I, I, P, P, P, P], None),
        "orc_tv_norm": ([I, I, I, I, P], D),
        "orc_residual_norm2": ([I, P, P, I64], D),
        "orc_tv_eps": ([I, I, I, I, P, D], D),
        "orc_tv_eps_grad": ([I, I, I, I, P, D, P], None),
        "orc_fdk_weight": ([G, I, P, P], None),
        "orc_ramp_filter": ([G, I, P, P, I, I], None),
        "orc_ramp_taps": ([G, I, P], None),
        "orc_fft": ([P, P, I, I], None),
        "orc_fdk_backproject": ([G, I, P, P, I, I], None),
        "orc_fdk": ([G, I, P, P, I], None),
        "orc_fdk_backproject_ms": ([G, I, P, P, I, I, I], None),
        "orc_fdk_ms": ([G, I, P, P, I, I], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_lib = _load()


def _np_dtype(prec: int):
    return np.float64 if prec == 64 else np.float32


def _arr(a: np.ndarray, prec: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=_np_dtype(prec))
    return a


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def set_threads(n: int) -> None:
    _lib.orc_set_threads(n)


def num_threads() -> int:
    return _lib.orc_num_threads()


def view_angle(g: Geometry, view: int, prec: int = 32, fdk: bool = False) -> tuple[float, float]:
    s, c = ctypes.c_double(), ctypes.c_double()
    gc = g.to_c()
    _lib.orc_view_angle(ctypes.byref(gc), prec, view, int(fdk), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def forward(g: Geometry, vol: np.ndarray, prec: int = 32, view0: int = 0, nviews: int | None = None) -> np.ndarray:
    nv = g.nviews - view0 if nviews is None else nviews
    v = _arr(vol, prec)
    out = np.zeros((nv, *g.proj_shape[1:]), dtype=_np_dtype(prec))
    gc = g.to_c()
    _lib.orc_forward(ctypes.byref(gc), prec, _p(v), _p(out), view0, nv)
    return out


def sart_correction(g: Geometry, vol: np.ndarray, meas: np.ndarray, view0: int, nviews: int = 1, prec: int = 32) -> np.ndarray:
    v, m = _arr(vol, prec), _arr(meas, prec)
    out = np.zeros((nviews, *g.proj_shape[1:]), dtype=_np_dtype(prec))
    gc = g.to_c()
    _lib.orc_sart_correction(ctypes.byref(gc), prec, _p(v), _p(m), _p(out), view0, nviews)
    return out


def backproject_voxel(g: Geometry, vol: np.ndarray, corr: np.ndarray, view0: int, nviews: int = 1,
                      beta: float = 1.0, mode: int = 1, prec: int = 32) -> np.ndarray:
    v, c = _arr(vol, prec).copy(), _arr(corr, prec)
    gc = g.to_c()
    _lib.orc_backproject_voxel(ctypes.byref(gc), prec, _p(v), _p(c), view0, nviews, beta, mode)
    return v


def sart_sweep(g: Geometry, vol: np.ndarray, meas: np.ndarray, beta: float = 1.0, subset: int = 1,
               order: str = "ascending", prec: int = 32) -> np.ndarray:
    v, m = _arr(vol, prec).copy(), _arr(meas, prec)
    gc = g.to_c()
    _lib.orc_sart_sweep(ctypes.byref(gc), prec, _p(v), _p(m), beta, subset, {"ascending": 0, "interleaved": 1}[order])
    return v


def adjoint(g: Geometry, proj: np.ndarray, prec: int = 64, view0: int = 0, nviews: int | None = None) -> np.ndarray:
    nv = g.nviews - view0 if nviews is None else nviews
    p = _arr(proj, prec)
    out = np.zeros(g.vol_shape, dtype=_np_dtype(prec))
    gc = g.to_c()
    _lib.orc_adjoint(ctypes.byref(gc), prec, _p(p), _p(out), view0, nv)
    return out


def ray_segments(g: Geometry, view: int, col: int, row: int, src: int = 0, prec: int = 64):
    cap = 4 * (g.nx + g.ny + g.nz) + 16
    idx = np.zeros(cap, dtype=np.int64)
    ln = np.zeros(cap, dtype=np.float64)
    gc = g.to_c()
    n = _lib.orc_ray_segments(ctypes.byref(gc), prec, view, col, row, src, _p(idx), _p(ln), cap)
    return idx[:n], ln[:n]


def fista_momentum(x: np.ndarray, xprev: np.ndarray, gamma: float, prec: int = 32):
    X, P = _arr(x, prec).copy(), _arr(xprev, prec).copy()
    Y = np.zeros_like(X)
    _lib.orc_fista_momentum(prec, _p(X), _p(P), _p(Y), X.size, gamma)
    return Y, P


def tv_fgp(b: np.ndarray, lam: float, niter: int = 20, prec: int = 32) -> np.ndarray:
    B = _arr(b, prec)
    out = np.zeros_like(B)
    nz, ny, nx = B.shape
    _lib.orc_tv_fgp(prec, nx, ny, nz, _p(B), _p(out), lam, niter)
    return out


def tv_fgp_duals(b: np.ndarray, lam: float, niter: int, prec: int = 64):
    """FGP result and its projected dual fields (o, p, q) after niter iterations."""
    B = _arr(b, prec)
    out = np.zeros_like(B)
    duals = np.zeros((3, *B.shape), dtype=B.dtype)
    nz, ny, nx = B.shape
    _lib.orc_tv_fgp_duals(prec, nx, ny, nz, _p(B), _p(out), _p(duals), lam, niter)
    return out, duals


def tv_D(x: np.ndarray, prec: int = 64):
    X = _arr(x, prec)
    nz, ny, nx = X.shape
    r, s, t = (np.zeros_like(X) for _ in range(3))
    _lib.orc_tv_D(prec, nx, ny, nz, _p(X), _p(r), _p(s), _p(t))
    return r, s, t


def tv_Dt(r: np.ndarray, s: np.ndarray, t: np.ndarray, prec: int = 64) -> np.ndarray:
    R, S, T = _arr(r, prec), _arr(s, prec), _arr(t, prec)
    nz, ny, nx = R.shape
    out = np.zeros_like(R)
    _lib.orc_tv_Dt(prec, nx, ny, nz, _p(R), _p(S), _p(T), _p(out))
    return out


def tv_norm(vol: np.ndarray, prec: int = 32) -> float:
    v = _arr(vol, prec)
    nz, ny, nx = v.shape
    return float(_lib.orc_tv_norm(prec, nx, ny, nz, _p(v)))


def residual_norm2(a: np.ndarray, b: np.ndarray, prec: int = 32) -> float:
    A, B = _arr(a, prec), _arr(b, prec)
    return float(_lib.orc_residual_norm2(prec, _p(A), _p(B), A.size))


def tv_eps(vol: np.ndarray, eps: float, prec: int = 64) -> float:
    v = _arr(vol, prec)
    nz, ny, nx = v.shape
    return float(_lib.orc_tv_eps(prec, nx, ny, nz, _p(v), eps))


def tv_eps_grad(vol: np.ndarray, eps: float, prec: int = 64) -> np.ndarray:
    v = _arr(vol, prec)
    nz, ny, nx = v.shape
    g = np.zeros_like(v)
    _lib.orc_tv_eps_grad(prec, nx, ny, nz, _p(v), eps, _p(g))
    return g


def fdk_weight(g: Geometry, proj: np.ndarray, prec: int = 32) -> np.ndarray:
    p = _arr(proj, prec)
    out = np.zeros_like(p)
    gc = g.to_c()
    _lib.orc_fdk_weight(ctypes.byref(gc), prec, _p(p), _p(out))
    return out


def ramp_filter(g: Geometry, proj: np.ndarray, window: str = "ramp", prec: int = 32) -> np.ndarray:
    p = _arr(proj, prec)
    out = np.zeros_like(p)
    nimg = p.size // (g.nrows * g.ncols)
    gc = g.to_c()
    _lib.orc_ramp_filter(ctypes.byref(gc), prec, _p(p), _p(out), nimg, {"ramp": 0, "hann": 1}[window])
    return out


def ramp_taps(g: Geometry, window: str = "ramp") -> np.ndarray:
    taps = np.zeros(2 * g.ncols, dtype=np.float64)
    gc = g.to_c()
    _lib.orc_ramp_taps(ctypes.byref(gc), {"ramp": 0, "hann": 1}[window], _p(taps))
    return taps


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    re = np.ascontiguousarray(np.real(x), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), dtype=np.float64).copy()
    _lib.orc_fft(_p(re), _p(im), re.size, sign)
    return re + 1j * im


def fdk_backproject(g: Geometry, filt: np.ndarray, vol: np.ndarray, view0: int = 0, nviews: int | None = None,
                    prec: int = 32) -> np.ndarray:
    nv = g.nviews - view0 if nviews is None else nviews
    f = _arr(filt, prec)
    v = _arr(vol, prec).copy()
    gc = g.to_c()
    _lib.orc_fdk_backproject(ctypes.byref(gc), prec, _p(f), _p(v), view0, nv)
    return v


FDK_SOURCES = {"inverse_square": 0, "nearest": 1, "equal": 2}


def fdk_backproject_ms(g: Geometry, filt: np.ndarray, vol: np.ndarray, mode: str = "inverse_square", view0: int = 0,
                       nviews: int | None = None, prec: int = 32) -> np.ndarray:
    """Multi-source FDK backprojection (any nsources, the TB kernels' path)."""
    nv = g.nviews - view0 if nviews is None else nviews
    f = _arr(filt, prec)
    v = _arr(vol, prec).copy()
    gc = g.to_c()
    _lib.orc_fdk_backproject_ms(ctypes.byref(gc), prec, _p(f), _p(v), view0, nv, FDK_SOURCES[mode])
    return v


def fdk(g: Geometry, proj: np.ndarray, window: str = "ramp", prec: int = 32, vol: np.ndarray | None = None,
        sources: str = "inverse_square") -> np.ndarray:
    p = _arr(proj, prec)
    v = np.zeros(g.vol_shape, dtype=_np_dtype(prec)) if vol is None else _arr(vol, prec).copy()
    gc = g.to_c()
    _lib.orc_fdk_ms(ctypes.byref(gc), prec, _p(p), _p(v), {"ramp": 0, "hann": 1}[window], FDK_SOURCES[sources])
    return v
