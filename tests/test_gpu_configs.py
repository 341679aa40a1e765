"""GPU parity at the benchmark configurations' own sizes and for the
reduced-precision path, through the C ABI, against the CPU oracle.

* 512^3 / 1024^2: FDK over 16 views and TV-FGP(20) at full size.
* 128^3 / 180 views / 256^2: the full ten SART iterations (subsets of 20).
* FP16 / BF16 storage with FP32 accumulation: SART sweeps and SART-FISTA-TV
  iterations against the FP32 oracle run on the storage-rounded inputs, and
  the 512^3 / 1024-view iteration against the GPU's FP64 instantiation.
* The SART step's power iteration against a dense eigen-decomposition.

Bars: plain FP32 path <= 1e-5 relative L2 on projections and <= 1e-4 on
reconstructions; FP16/BF16 paths are held to the RMSEs stated per test.
"""
import math
from dataclasses import replace

import numpy as np
import pytest
import torch

from gpu_amd.geometry import Geometry, plain_sart_128, repo_scale_512
from gpu_amd.phantom import phantom, poisson_log
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rmse(a, b):
    d = np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)
    return float(np.sqrt((d * d).mean()))


def dev(x, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype).contiguous()


def host(t):
    torch.cuda.synchronize()
    return t.detach().to("cpu", torch.float64).numpy()


def small_geom(**kw):
    base = dict(nx=24, ny=20, nz=16, vx=1.0, ncols=40, nrows=28, du=1.1, nviews=9, dso=90.0, dod=-45.0,
                angle_step_deg=37.3)
    base.update(kw)
    return Geometry(**base)


def test_fdk_full_size_16_views(ctx):
    """FDK with the Hann window (V63/host_FDK.cpp:4-57 plus the window) at the
    benchmark size: 512^3 at 0.5 mm from 16 views of 1024^2 spread over 360
    degrees: cosine weight fused into the radix-8 filter (pad 2048), then the
    LDS-staged backprojector, against the FP32 restatement at the 1e-4
    reconstruction bar."""
    g = replace(repo_scale_512(), nviews=16)
    x = phantom(g, seed=0, scale=2.5, device="cuda")
    proj = ctx.forward(g, x)
    p_host = host(proj).astype(np.float32)
    ref = orc.fdk(g, p_host, "hann", prec=32)
    got = host(ctx.fdk(g, proj, window="hann"))
    r = rel(got, ref)
    print(f"FDK 512^3 from 16 views: rel L2 {r:.3e}")
    assert r <= 1e-4


def test_tv_fgp_full_size(ctx):
    """TV-FGP with 20 inner iterations and lambda 1e-3 (V63/host_FGP_Denoise_GPUx7.c:64-86)
    on the noisy 512^3 phantom: the fused step over its real tile and z-chunk
    counts, against the FP32 restatement to 1e-5."""
    g = repo_scale_512()
    x = phantom(g, seed=0, scale=2.5, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(3)
    b = (x + 0.002 * torch.randn(x.shape, device="cuda", generator=gen)).contiguous()
    ref = orc.tv_fgp(host(b).astype(np.float32), 1e-3, 20, prec=32)
    ctx.tv_fgp(b, 1e-3, 20)
    r = rel(host(b), ref)
    print(f"TV-FGP(20) 512^3: rel L2 {r:.3e}")
    assert r <= 1e-5


def test_plain_sart_128_ten_iterations(ctx):
    """The CPU-runnable configuration in full: 128^3 at 1 mm, 180 views of
    256^2, SART with relaxation 1.0 in interleaved subsets of 20 views, ten
    iterations from zero, FP32, against the FP32 oracle at the 1e-4 bar; the
    error against the FP64 oracle and the phantom is reported."""
    g = plain_sart_128()
    x = phantom(g, seed=0).numpy()
    b = orc.forward(g, x, prec=32)
    v = torch.zeros(g.vol_shape, device="cuda")
    bd = dev(b)
    ref = np.zeros_like(x)
    ref64 = np.zeros(x.shape)
    b64 = orc.forward(g, x.astype(np.float64), prec=64)
    for _ in range(10):
        ctx.sart_sweep(g, v, bd, beta=1.0, subset=20, order="interleaved")
        ref = orc.sart_sweep(g, ref, b, beta=1.0, subset=20, order="interleaved", prec=32)
        ref64 = orc.sart_sweep(g, ref64, b64, beta=1.0, subset=20, order="interleaved", prec=64)
    got = host(v)
    r = rel(got, ref)
    print(f"plain SART 128^3 x 10: vs FP32 oracle {r:.3e}, vs FP64 oracle {rel(got, ref64):.3e} "
          f"(FP32 oracle vs FP64 {rel(ref, ref64):.3e}), vs phantom {rel(got, x):.3e}")
    assert r <= 1e-4


RP = {torch.float16: ("f16", 3e-3), torch.bfloat16: ("bf16", 2e-2)}


@pytest.mark.parametrize("dt", [torch.float16, torch.bfloat16])
@pytest.mark.parametrize("name,kw,subset", [("small", {}, 1), ("offsets", dict(x_offset=3, det_offset=2.2), 3),
                                            ("multisource", dict(nsources=4, src_z0=-9.0, src_dz=6.0,
                                                                 nrows=12), 1)])
def test_reduced_precision_sweep(ctx, dt, name, kw, subset):
    """Two SART sweeps on FP16 / BF16 storage (z-quad copy, packed-half
    chord engine, stochastically rounded updates) against the FP32 oracle
    started from the same storage-rounded volume and measurements.  Each
    view's update is rounded to the storage type once, so the gap is a few
    storage ulps: relative L2 <= 3e-3 (FP16) and 2e-2 (BF16)."""
    g = small_geom(**kw)
    x = phantom(g, seed=15, scale=0.3).numpy()
    b = orc.forward(g, x, prec=32)
    v0 = torch.as_tensor(x * 0.5).to(dt)
    bd = dev(b).to(dt)
    ref = v0.float().numpy()
    bref = bd.float().cpu().numpy()
    v = v0.cuda()
    for _ in range(2):
        ctx.sart_sweep(g, v, bd, beta=0.9, subset=subset)
        ref = orc.sart_sweep(g, ref, bref, beta=0.9, subset=subset, prec=32)
    got = host(v)
    tag, tol = RP[dt]
    r = rel(got, ref)
    print(f"{name} {tag}: sweep rel L2 {r:.3e}, RMSE {rmse(got, ref):.3e}")
    assert r <= tol


@pytest.mark.parametrize("dt", [torch.float16, torch.bfloat16])
def test_reduced_precision_iteration_128(ctx, dt):
    """Three SART-FISTA-TV iterations (V63/main_recon.cu:344-385) at 128^3 /
    180 views (subsets of 20, relaxation 1, FGP(10, 1e-3)) on FP16 / BF16
    storage against the FP32 oracle's composition of sweep, FGP and momentum
    from the storage-rounded measurements: RMSE <= 2e-4 /mm (FP16) and
    1e-3 /mm (BF16) on volumes of ~0.02 /mm."""
    g = plain_sart_128()
    x = phantom(g, seed=0).numpy()
    b = orc.forward(g, x, prec=32)
    bd = dev(b).to(dt)
    bref = bd.float().cpu().numpy()
    X = torch.zeros(g.vol_shape, device="cuda", dtype=dt)
    XP = torch.zeros_like(X)
    y_ref, xp_ref, t, t_gpu = np.zeros_like(x), np.zeros_like(x), 1.0, 1.0
    for _ in range(3):
        t_gpu = ctx.sart_fista_iteration(g, X, XP, bd, t_gpu, beta=1.0, subset=20, order="interleaved",
                                         fgp_lambda=1e-3, fgp_niter=10)
        xk = orc.sart_sweep(g, y_ref, bref, beta=1.0, subset=20, order="interleaved", prec=32)
        xk = orc.tv_fgp(xk, 1e-3, 10, prec=32)
        tn = (1 + math.sqrt(1 + 4 * t * t)) / 2
        y_ref, xp_ref = orc.fista_momentum(xk, xp_ref, (t - 1) / tn, prec=32)
        t = tn
    got = host(XP)
    e, r = rmse(got, xp_ref), rel(got, xp_ref)
    tag = RP[dt][0]
    print(f"128^3 SART-FISTA-TV x 3, {tag}: RMSE {e:.3e} /mm, rel L2 {r:.3e}")
    assert t_gpu == pytest.approx(t)
    assert e <= (2e-4 if dt == torch.float16 else 1e-3)


def test_reduced_precision_iteration_512(ctx):
    """The reduced-precision configuration at its own size: 512^3 / 1024 views
    of 1024^2, Poisson-noise measurements, SART step from the power
    iteration of the view-by-view sweep operator (2/Lip ~ 1.14e-3), TV-FGP(20,
    1e-3) and FISTA momentum; two iterations each in FP64 (the oracle
    instantiation on the GPU), FP32, FP16 and BF16 storage.  RMSE against
    FP64 within 1e-3 /mm for both 16-bit types (the target bar); FP32 within
    1e-5 /mm (its rounding, carried through 2048 view updates)."""
    g = repo_scale_512()
    x_true = phantom(g, seed=0, scale=2.5, device="cuda")
    meas = poisson_log(ctx.forward(g, x_true), i0=1e5, seed=1)
    del x_true
    lip = ctx.sart_lipschitz(g, torch.float32, niter=30, seed=1, subset=1)
    beta = 2.0 / lip
    print(f"SART power iteration: lipschitz {lip:.6f}, step {beta:.6e}")
    # each view's term has radius <= 1, so the sweep's is <= 1024 views
    assert 1024 < lip <= 2 * 1024
    out = {}
    for dt in (torch.float64, torch.float32, torch.float16, torch.bfloat16):
        m = meas if dt in (torch.float32, torch.float64) else meas.to(dt)
        X = torch.zeros(g.vol_shape, device="cuda", dtype=dt)
        XP = torch.zeros_like(X)
        t = 1.0
        for _ in range(2):
            t = ctx.sart_fista_iteration(g, X, XP, m, t, beta=beta, subset=1, fgp_lambda=1e-3, fgp_niter=20)
        torch.cuda.synchronize()
        out[dt] = XP.double()
        del X, m
    ref = out[torch.float64]
    for dt, bar in ((torch.float32, 1e-5), (torch.float16, 1e-3), (torch.bfloat16, 1e-3)):
        d = out[dt] - ref
        e = float(torch.sqrt((d * d).mean()))
        r = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref))
        print(f"512^3 x 2 iterations, {dt}: RMSE {e:.3e} /mm, rel L2 {r:.3e}")
        assert e <= bar


@pytest.mark.parametrize("subset", [1, 2, 5])
def test_sart_lipschitz_vs_dense(ctx, subset):
    """ct_sart_lipschitz against the dense operator of one SART sweep with
    subsets of `subset` views, K = sum_S D_S^-1 B_S^T M_S^-1 A_S, built column
    by column from the oracle's correction (zero measurements) and
    voxel-driven backprojection of each subset: the power iteration's radius
    equals the spectral radius to 1e-3 (400 iterations).  K has non-negative
    entries, so that radius is a real eigenvalue with a non-negative
    eigenvector (Perron-Frobenius).  Each subset's term averages along rays and
    over a voxel's rays, so its radius is at most 1 and the sweep's at most
    the number of subsets; the simultaneous operator (subset 5 = all views)
    has radius at most 1 (0.59 on this 6 x 6 x 4 grid: the reconstruction
    mask and the rays that leave through masked voxels lower it)."""
    g = Geometry(nx=6, ny=6, nz=4, vx=1.0, ncols=9, nrows=6, du=1.2, nviews=5, dso=40.0, dod=-20.0)
    n = g.nvox
    K = np.zeros((n, n))
    zero = np.zeros(g.proj_shape)
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        for p0 in range(0, g.nviews, subset):
            cnt = min(subset, g.nviews - p0)
            c = orc.sart_correction(g, e.reshape(g.vol_shape), zero, view0=p0, nviews=cnt, prec=64)
            K[:, k] += orc.backproject_voxel(g, np.zeros(g.vol_shape), c, view0=p0, nviews=cnt, beta=-1.0,
                                             mode=1, prec=64).ravel()
    assert K.min() >= 0.0
    ev = np.linalg.eigvals(K)
    rho = float(np.abs(ev).max())
    assert float(ev.real.max()) == pytest.approx(rho, rel=1e-12)
    lip = ctx.sart_lipschitz(g, torch.float64, niter=400, seed=3, subset=subset)
    nsub = -(-g.nviews // subset)
    print(f"subset {subset}: dense radius {rho:.9f}, power iteration {lip / 2:.9f}")
    assert rho <= nsub * (1.0 + 1e-9)
    assert lip / 2 == pytest.approx(rho, rel=1e-3)
