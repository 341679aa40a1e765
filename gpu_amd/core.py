"""Python binding of the C ABI (include/ct_core.h) over torch device tensors.

torch only supplies device memory and the current HIP stream; every operation
runs in the HIP kernels of libctcore.so.  There is no CPU fallback: importing
this module without the built library raises.

Method names follow the reference host functions they replace, e.g.
``forward`` = Forward_3D_ray_driven_siddon
(CBCT/CBCT_1024V_1024P_multiGPU_v6.3_20170823/host_IterativeRecon_CBCT.c:70),
``sart_sweep`` = Reconstruction_3D_ray_driven_CBCT (same file :1),
``tv_fgp`` = FGP_denoise_GPUx7_exact (host_FGP_Denoise_GPUx7.c:1),
``fdk`` = Reconstruction_FDK (host_FDK.cpp:4).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import torch

from .geometry import CtGeometry, Geometry

LIB_PATH = Path(__file__).resolve().parent / "libctcore.so"

F32, F16, BF16, F64 = 0, 1, 2, 3
_DT = {torch.float32: F32, torch.float16: F16, torch.bfloat16: BF16, torch.float64: F64}
ORDER = {"ascending": 0, "interleaved": 1}
KERNELS = ("forward", "voxel_bp", "adjoint", "fgp", "fista", "fdk_bp", "ramp")
WINDOW = {"ramp": 0, "hann": 1}
FDK_SOURCES = {"inverse_square": 0, "nearest": 1, "equal": 2}


class CtError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ct status {status}: {message}")
        self.status = status


def _load() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m gpu_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(str(LIB_PATH))
    P, G, I, D, U64, DP = ctypes.c_void_p, ctypes.POINTER(CtGeometry), ctypes.c_int32, ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)
    I64 = ctypes.c_int64
    sig = {
        "ct_abi_version": ([], ctypes.c_int),
        "ct_create": ([ctypes.POINTER(P), ctypes.c_int], I),
        "ct_destroy": ([P], I),
        "ct_set_stream": ([P, P], I),
        "ct_synchronize": ([P], I),
        "ct_last_error": ([P], ctypes.c_char_p),
        "ct_status_string": ([I], ctypes.c_char_p),
        "ct_validate_geometry": ([G], I),
        "ct_view_angle": ([G, I, I, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)], I),
        "ct_scratch_bytes": ([P], ctypes.c_size_t),
        "ct_forward": ([P, G, P, I, P, I, I, I], I),
        "ct_sart_correction": ([P, G, P, I, P, I, P, I, I], I),
        "ct_backproject_voxel": ([P, G, P, I, P, I, I, D, I], I),
        "ct_sart_sweep": ([P, G, P, I, P, I, D, I, I], I),
        "ct_adjoint": ([P, G, P, I, P, I, I, I], I),
        "ct_adjoint_test": ([P, G, I, U64, DP, DP], I),
        "ct_power_iteration": ([P, G, I, I, U64, DP], I),
        "ct_sart_lipschitz": ([P, G, I, I, I, U64, DP], I),
        "ct_fista_momentum": ([P, P, P, P, I64, I, D], I),
        "ct_fista_gamma": ([D, DP, DP], I),
        "ct_tv_fgp": ([P, I, I, I, P, I, D, I], I),
        "ct_tv_fgp_adjoint": ([P, I, I, I, P, P, P, P, I], I),
        "ct_tv_fgp_forward_op": ([P, I, I, I, P, P, P, P, I], I),
        "ct_tv_eps": ([P, I, I, I, P, I, D, DP], I),
        "ct_tv_eps_grad": ([P, I, I, I, P, I, D, P], I),
        "ct_residual_norm2": ([P, P, P, I64, I, DP], I),
        "ct_tv_norm": ([P, I, I, I, P, I, DP], I),
        "ct_dot": ([P, P, P, I64, I, DP], I),
        "ct_fdk_weight": ([P, G, P, P, I], I),
        "ct_ramp_filter": ([P, G, P, P, I, I, I], I),
        "ct_fdk_backproject": ([P, G, P, I, P, I, I, I], I),
        "ct_fdk": ([P, G, P, I, P, I, I], I),
        "ct_fdk_ms": ([P, G, P, I, P, I, I, I], I),
        "ct_fdk_backproject_ms": ([P, G, P, I, P, I, I, I, I], I),
        "ct_sart_fista_iteration": ([P, G, P, P, I, P, I, D, I, I, D, I, DP], I),
        "ct_convert": ([P, P, I, P, I, I64], I),
        "ct_objective": ([P, G, P, I, P, I, D, DP, DP, DP], I),
        "ct_set_math": ([P, I], I),
        "ct_profile_enable": ([P, I], I),
        "ct_profile_reset": ([P], I),
        "ct_profile_read": ([P, I, DP, ctypes.POINTER(ctypes.c_int64)], I),
        "ct_projection_path": ([G, I, I, ctypes.c_char_p, ctypes.c_char_p, I], I),
        "ct_read_projections": ([P, G, ctypes.c_char_p, I, I, P, I, I, I], I),
        "ct_log_transform": ([P, P, I, P, P, D, P, I, I64, I, D], I),
        "ct_write_volume": ([P, G, P, I, ctypes.c_char_p], I),
        "ct_read_volume": ([P, G, ctypes.c_char_p, P, I], I),
        "ct_recon_path": ([G, D, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, I, D, D, D, ctypes.c_char_p, I], I),
        "ct_objective_path": ([G, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, D, D, ctypes.c_char_p, I], I),
        "ct_write_objective": ([ctypes.c_char_p, DP, I64], I),
        "ct_save_checkpoint": ([P, G, ctypes.c_char_p, P, P, I, D, I], I),
        "ct_load_checkpoint": ([P, G, ctypes.c_char_p, P, P, I, DP, ctypes.POINTER(ctypes.c_int32)], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_lib = _load()


def lib() -> ctypes.CDLL:
    return _lib


def exported_symbols() -> list[str]:
    return [n for n in dir(_lib) if n.startswith("ct_")]


def dtype_code(t: torch.Tensor) -> int:
    if t.dtype not in _DT:
        raise CtError(3, f"unsupported tensor dtype {t.dtype}")
    return _DT[t.dtype]


def _ptr(t: torch.Tensor | None):
    if t is None:
        return None
    if not t.is_cuda:
        raise CtError(1, "tensor must live on the GPU")
    if not t.is_contiguous():
        raise CtError(1, "tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _need(t: torch.Tensor, numel: int, name: str) -> None:
    if t.numel() != numel:
        raise CtError(1, f"{name} has {t.numel()} elements, expected {numel}")


class Context:
    """A ct_context on one GPU; calls run on torch's current stream."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        st = _lib.ct_create(ctypes.byref(self._h), device)
        if st != 0:
            raise CtError(st, _lib.ct_status_string(st).decode())
        self.device = device

    def close(self) -> None:
        if self._h:
            _lib.ct_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> None:
        if st != 0:
            raise CtError(st, _lib.ct_last_error(self._h).decode())

    def _stream(self) -> None:
        _lib.ct_set_stream(self._h, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))

    def synchronize(self) -> None:
        self._check(_lib.ct_synchronize(self._h))

    @property
    def scratch_bytes(self) -> int:
        return int(_lib.ct_scratch_bytes(self._h))

    def set_math(self, mode: str) -> None:
        """'exact': the reference's FP32 operation sequence (bit-identical to the
        CPU restatement); 'fast' (default): hardware reciprocal/sqrt and FMA in
        the voxel-driven backprojector."""
        self._check(_lib.ct_set_math(self._h, {"exact": 0, "fast": 1}[mode]))

    # ---- kernel timing (HIP events on the context stream)
    def profile(self, enable: bool = True, reset: bool = True) -> None:
        if reset:
            self._check(_lib.ct_profile_reset(self._h))
        self._check(_lib.ct_profile_enable(self._h, int(enable)))

    def profile_read(self) -> dict[str, tuple[float, int]]:
        out = {}
        for k, name in enumerate(KERNELS):
            ms, n = ctypes.c_double(), ctypes.c_int64()
            self._check(_lib.ct_profile_read(self._h, k, ctypes.byref(ms), ctypes.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    # ---- projector pair
    def forward(self, g: Geometry, vol: torch.Tensor, proj: torch.Tensor | None = None,
                view0: int = 0, nviews: int | None = None, out_dtype: torch.dtype | None = None) -> torch.Tensor:
        nv = g.nviews - view0 if nviews is None else nviews
        if proj is None:
            proj = torch.empty((nv, *g.proj_shape[1:]), device=vol.device, dtype=out_dtype or vol.dtype)
        _need(vol, g.nvox, "vol")
        _need(proj, nv * g.nsources * g.nrows * g.ncols, "proj")
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_forward(self._h, ctypes.byref(gc), _ptr(vol), dtype_code(vol), _ptr(proj), dtype_code(proj), view0, nv))
        return proj

    def sart_correction(self, g: Geometry, vol: torch.Tensor, meas: torch.Tensor, view0: int = 0,
                        nviews: int = 1, corr: torch.Tensor | None = None) -> torch.Tensor:
        cdt = torch.float64 if vol.dtype == torch.float64 else torch.float32
        if corr is None:
            corr = torch.empty((nviews, *g.proj_shape[1:]), device=vol.device, dtype=cdt)
        _need(meas, nviews * g.nsources * g.nrows * g.ncols, "meas")
        _need(corr, nviews * g.nsources * g.nrows * g.ncols, "corr")
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_sart_correction(self._h, ctypes.byref(gc), _ptr(vol), dtype_code(vol), _ptr(meas),
                                            dtype_code(meas), _ptr(corr), view0, nviews))
        return corr

    def backproject_voxel(self, g: Geometry, vol: torch.Tensor, corr: torch.Tensor, view0: int = 0,
                          nviews: int = 1, beta: float = 1.0, mode: int = 1) -> torch.Tensor:
        _need(corr, nviews * g.nsources * g.nrows * g.ncols, "corr")
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_backproject_voxel(self._h, ctypes.byref(gc), _ptr(vol), dtype_code(vol), _ptr(corr),
                                              view0, nviews, beta, mode))
        return vol

    def sart_sweep(self, g: Geometry, vol: torch.Tensor, meas: torch.Tensor, beta: float = 1.0,
                   subset: int = 1, order: str = "ascending") -> torch.Tensor:
        _need(vol, g.nvox, "vol")
        _need(meas, g.rays, "meas")
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_sart_sweep(self._h, ctypes.byref(gc), _ptr(vol), dtype_code(vol), _ptr(meas),
                                       dtype_code(meas), beta, subset, ORDER[order]))
        return vol

    def adjoint(self, g: Geometry, proj: torch.Tensor, vol: torch.Tensor | None = None, view0: int = 0,
                nviews: int | None = None) -> torch.Tensor:
        nv = g.nviews - view0 if nviews is None else nviews
        if vol is None:
            vol = torch.empty(g.vol_shape, device=proj.device, dtype=proj.dtype)
        _need(proj, nv * g.nsources * g.nrows * g.ncols, "proj")
        _need(vol, g.nvox, "vol")
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_adjoint(self._h, ctypes.byref(gc), _ptr(proj), dtype_code(proj), _ptr(vol),
                                    dtype_code(vol), view0, nv))
        return vol

    def adjoint_test(self, g: Geometry, dtype: torch.dtype = torch.float64, seed: int = 0) -> tuple[float, float]:
        lhs, rhs = ctypes.c_double(), ctypes.c_double()
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_adjoint_test(self._h, ctypes.byref(gc), _DT[dtype], seed, ctypes.byref(lhs), ctypes.byref(rhs)))
        return lhs.value, rhs.value

    def power_iteration(self, g: Geometry, dtype: torch.dtype = torch.float32, niter: int = 30, seed: int = 0) -> float:
        lip = ctypes.c_double()
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_power_iteration(self._h, ctypes.byref(gc), _DT[dtype], niter, seed, ctypes.byref(lip)))
        return lip.value

    def sart_lipschitz(self, g: Geometry, dtype: torch.dtype = torch.float32, niter: int = 30, seed: int = 0,
                       subset: int = 1) -> float:
        """2 x the largest eigenvalue of the operator one SART sweep with
        subsets of `subset` views applies to first order in its step
        (ct_sart_lipschitz); the SART-FISTA step is 2 / lipschitz."""
        lip = ctypes.c_double()
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_sart_lipschitz(self._h, ctypes.byref(gc), _DT[dtype], subset, niter, seed,
                                           ctypes.byref(lip)))
        return lip.value

    # ---- FISTA / TV
    def fista_momentum(self, x: torch.Tensor, x_prev: torch.Tensor, y: torch.Tensor, gamma: float) -> None:
        self._stream()
        self._check(_lib.ct_fista_momentum(self._h, _ptr(x), _ptr(x_prev), _ptr(y), x.numel(), dtype_code(x), gamma))

    @staticmethod
    def fista_gamma(t_prev: float) -> tuple[float, float]:
        tn, gm = ctypes.c_double(), ctypes.c_double()
        st = _lib.ct_fista_gamma(t_prev, ctypes.byref(tn), ctypes.byref(gm))
        if st != 0:
            raise CtError(st, "invalid t_prev")
        return tn.value, gm.value

    def tv_fgp(self, vol: torch.Tensor, lam: float, niter: int = 20) -> torch.Tensor:
        nz, ny, nx = vol.shape
        self._stream()
        self._check(_lib.ct_tv_fgp(self._h, nx, ny, nz, _ptr(vol), dtype_code(vol), lam, niter))
        return vol

    def tv_fgp_D(self, x: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        nz, ny, nx = x.shape
        r, s, t = (torch.empty_like(x) for _ in range(3))
        self._stream()
        self._check(_lib.ct_tv_fgp_forward_op(self._h, nx, ny, nz, _ptr(x), _ptr(r), _ptr(s), _ptr(t), dtype_code(x)))
        return r, s, t

    def tv_fgp_Dt(self, r: torch.Tensor, s: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
        nz, ny, nx = r.shape
        out = torch.empty_like(r)
        self._stream()
        self._check(_lib.ct_tv_fgp_adjoint(self._h, nx, ny, nz, _ptr(r), _ptr(s), _ptr(t), _ptr(out), dtype_code(r)))
        return out

    def tv_eps(self, vol: torch.Tensor, eps: float) -> float:
        nz, ny, nx = vol.shape
        v = ctypes.c_double()
        self._stream()
        self._check(_lib.ct_tv_eps(self._h, nx, ny, nz, _ptr(vol), dtype_code(vol), eps, ctypes.byref(v)))
        return v.value

    def tv_eps_grad(self, vol: torch.Tensor, eps: float, grad: torch.Tensor | None = None) -> torch.Tensor:
        nz, ny, nx = vol.shape
        grad = torch.empty_like(vol) if grad is None else grad
        self._stream()
        self._check(_lib.ct_tv_eps_grad(self._h, nx, ny, nz, _ptr(vol), dtype_code(vol), eps, _ptr(grad)))
        return grad

    def residual_norm2(self, a: torch.Tensor, b: torch.Tensor) -> float:
        v = ctypes.c_double()
        self._stream()
        self._check(_lib.ct_residual_norm2(self._h, _ptr(a), _ptr(b), a.numel(), dtype_code(a), ctypes.byref(v)))
        return v.value

    def tv_norm(self, vol: torch.Tensor) -> float:
        nz, ny, nx = vol.shape
        v = ctypes.c_double()
        self._stream()
        self._check(_lib.ct_tv_norm(self._h, nx, ny, nz, _ptr(vol), dtype_code(vol), ctypes.byref(v)))
        return v.value

    def objective(self, g: Geometry, vol: torch.Tensor, meas: torch.Tensor, lambda_tv: float = 0.0
                  ) -> tuple[float, float, float]:
        """(sum (Ax-b)^2, TV(x), data + 2*lambda_tv*TV) as V63/main_recon.cu:401-428 forms them."""
        _need(meas, g.rays, "meas")
        d, t, o = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_objective(self._h, ctypes.byref(gc), _ptr(vol), dtype_code(vol), _ptr(meas),
                                      dtype_code(meas), lambda_tv, ctypes.byref(d), ctypes.byref(t), ctypes.byref(o)))
        return d.value, t.value, o.value

    def dot(self, a: torch.Tensor, b: torch.Tensor) -> float:
        v = ctypes.c_double()
        self._stream()
        self._check(_lib.ct_dot(self._h, _ptr(a), _ptr(b), a.numel(), dtype_code(a), ctypes.byref(v)))
        return v.value

    # ---- FDK
    def fdk_weight(self, g: Geometry, proj: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        out = torch.empty_like(proj) if out is None else out
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_fdk_weight(self._h, ctypes.byref(gc), _ptr(proj), _ptr(out), dtype_code(proj)))
        return out

    def ramp_filter(self, g: Geometry, proj: torch.Tensor, out: torch.Tensor | None = None,
                    window: str = "ramp") -> torch.Tensor:
        out = torch.empty_like(proj) if out is None else out
        nimg = proj.numel() // (g.nrows * g.ncols)
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_ramp_filter(self._h, ctypes.byref(gc), _ptr(proj), _ptr(out), dtype_code(proj), nimg, WINDOW[window]))
        return out

    def fdk_backproject(self, g: Geometry, filt: torch.Tensor, vol: torch.Tensor, view0: int = 0,
                        nviews: int | None = None, sources: str = "inverse_square") -> torch.Tensor:
        """Multi-source geometries combine their sources per view as `sources` says
        (TB/FDK: inverse_square = kernel_FDK.cu, nearest = OneRay, equal = EqualWeighting)."""
        nv = g.nviews - view0 if nviews is None else nviews
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_fdk_backproject_ms(self._h, ctypes.byref(gc), _ptr(filt), dtype_code(filt), _ptr(vol),
                                               dtype_code(vol), view0, nv, FDK_SOURCES[sources]))
        return vol

    def fdk(self, g: Geometry, proj: torch.Tensor, vol: torch.Tensor | None = None, window: str = "ramp",
            sources: str = "inverse_square") -> torch.Tensor:
        if vol is None:
            vol = torch.zeros(g.vol_shape, device=proj.device, dtype=proj.dtype)
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_fdk_ms(self._h, ctypes.byref(gc), _ptr(proj), dtype_code(proj), _ptr(vol),
                                   dtype_code(vol), WINDOW[window], FDK_SOURCES[sources]))
        return vol

    # ---- driver
    def sart_fista_iteration(self, g: Geometry, x: torch.Tensor, x_prev: torch.Tensor, meas: torch.Tensor,
                             t_state: float, beta: float = 1.0, subset: int = 1, order: str = "ascending",
                             fgp_lambda: float = 0.0, fgp_niter: int = 20) -> float:
        t = ctypes.c_double(t_state)
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_sart_fista_iteration(self._h, ctypes.byref(gc), _ptr(x), _ptr(x_prev), dtype_code(x),
                                                 _ptr(meas), dtype_code(meas), beta, subset, ORDER[order],
                                                 fgp_lambda, fgp_niter, ctypes.byref(t)))
        return t.value

    def convert(self, src: torch.Tensor, dst: torch.Tensor) -> torch.Tensor:
        self._stream()
        self._check(_lib.ct_convert(self._h, _ptr(src), dtype_code(src), _ptr(dst), dtype_code(dst), src.numel()))
        return dst

    # ---- files (main_recon.cu:196-224 projections in, :437-487 outputs)
    def read_projections(self, g: Geometry, directory: str, proj: torch.Tensor | None = None, view0: int = 0,
                         nviews: int | None = None, row0: int = 0, tomo: bool = False,
                         dtype: torch.dtype = torch.float32) -> torch.Tensor:
        nv = g.nviews - view0 if nviews is None else nviews
        if proj is None:
            proj = torch.empty((nv, g.nrows, g.ncols), device="cuda", dtype=dtype)
        _need(proj, nv * g.nrows * g.ncols, "proj")
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_read_projections(self._h, ctypes.byref(gc), directory.encode(), int(tomo), row0,
                                             _ptr(proj), dtype_code(proj), view0, nv))
        return proj

    def log_transform(self, raw: torch.Tensor, flat: torch.Tensor | None = None, dark: torch.Tensor | None = None,
                      i0: float = 1.0, out: torch.Tensor | None = None, floor: float = 1e-6) -> torch.Tensor:
        """-ln(max((raw - dark)/(flat - dark), floor)); flat and dark are one image, broadcast over raw's
        leading dimension (flat None: the scalar i0, dark None: 0)."""
        pix = raw.shape[-2] * raw.shape[-1]
        out = torch.empty_like(raw, dtype=torch.float32) if out is None else out
        for t, name in ((flat, "flat"), (dark, "dark")):
            if t is not None:
                if t.dtype != torch.float32:
                    raise CtError(3, f"{name} must be float32")
                _need(t, pix, name)
        self._stream()
        self._check(_lib.ct_log_transform(self._h, _ptr(raw), dtype_code(raw), _ptr(flat), _ptr(dark), i0, _ptr(out),
                                          dtype_code(out), pix, raw.numel() // pix, floor))
        return out

    def write_volume(self, g: Geometry, vol: torch.Tensor, path: str) -> None:
        _need(vol, g.nx * g.ny * g.nz, "vol")
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_write_volume(self._h, ctypes.byref(gc), _ptr(vol), dtype_code(vol), str(path).encode()))

    def read_volume(self, g: Geometry, path: str, vol: torch.Tensor | None = None,
                    dtype: torch.dtype = torch.float32) -> torch.Tensor:
        vol = torch.empty(g.vol_shape, device="cuda", dtype=dtype) if vol is None else vol
        _need(vol, g.nx * g.ny * g.nz, "vol")
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_read_volume(self._h, ctypes.byref(gc), str(path).encode(), _ptr(vol), dtype_code(vol)))
        return vol

    def save_checkpoint(self, g: Geometry, path: str, x: torch.Tensor, x_prev: torch.Tensor, t_state: float,
                        iteration: int) -> None:
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_save_checkpoint(self._h, ctypes.byref(gc), str(path).encode(), _ptr(x), _ptr(x_prev),
                                            dtype_code(x), t_state, iteration))

    def load_checkpoint(self, g: Geometry, path: str, x: torch.Tensor, x_prev: torch.Tensor) -> tuple[float, int]:
        t, k = ctypes.c_double(), ctypes.c_int32()
        self._stream()
        gc = g.to_c()
        self._check(_lib.ct_load_checkpoint(self._h, ctypes.byref(gc), str(path).encode(), _ptr(x), _ptr(x_prev),
                                            dtype_code(x), ctypes.byref(t), ctypes.byref(k)))
        return t.value, k.value


def validate_geometry(g: Geometry) -> int:
    gc = g.to_c()
    return int(_lib.ct_validate_geometry(ctypes.byref(gc)))


def view_angle(g: Geometry, view: int, fdk: bool = False) -> tuple[float, float]:
    s, c = ctypes.c_float(), ctypes.c_float()
    gc = g.to_c()
    st = _lib.ct_view_angle(ctypes.byref(gc), view, int(fdk), ctypes.byref(s), ctypes.byref(c))
    if st != 0:
        raise CtError(st, _lib.ct_status_string(st).decode())
    return s.value, c.value


def _path(st: int, buf) -> str:
    if st != 0:
        raise CtError(st, _lib.ct_status_string(st).decode())
    return buf.value.decode()


def projection_path(g: Geometry, view: int, directory: str, tomo: bool = False) -> str:
    """The reference's per-view file name (main_recon.cu:199-224)."""
    buf = ctypes.create_string_buffer(4096)
    gc = g.to_c()
    return _path(_lib.ct_projection_path(ctypes.byref(gc), view, int(tomo), directory.encode(), buf, 4096), buf)


def recon_path(g: Geometry, directory: str, output_folder: str, object_name: str, iteration: int = -1,
               lambda_tv: float = 0.0, lambda_l1: float = 0.0, lipschitz: float = 32.0, unit_um: float = 1e3) -> str:
    """The reference's .recon name (main_recon.cu:438,452): iteration < 0 gives the running file."""
    buf = ctypes.create_string_buffer(4096)
    gc = g.to_c()
    return _path(_lib.ct_recon_path(ctypes.byref(gc), unit_um, directory.encode(), output_folder.encode(),
                                    object_name.encode(), iteration, lambda_tv, lambda_l1, lipschitz, buf, 4096), buf)


def objective_path(g: Geometry, directory: str, output_folder: str, object_name: str, lambda_tv: float = 0.0,
                   lipschitz: float = 32.0) -> str:
    """The reference's object_func file name (main_recon.cu:468)."""
    buf = ctypes.create_string_buffer(4096)
    gc = g.to_c()
    return _path(_lib.ct_objective_path(ctypes.byref(gc), directory.encode(), output_folder.encode(),
                                        object_name.encode(), lambda_tv, lipschitz, buf, 4096), buf)


def write_objective(path: str, values) -> None:
    """(tv, data fidelity, objective) triples as doubles (main_recon.cu:467-477)."""
    arr = (ctypes.c_double * len(values))(*[float(v) for v in values])
    st = _lib.ct_write_objective(str(path).encode(), arr, len(values))
    if st != 0:
        raise CtError(st, _lib.ct_status_string(st).decode())
