"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Bars: plain FP32 path <= 1e-5 relative L2 on
projections and <= 1e-4 on reconstructions; FP64 paths are held to 1e-12;
the matched adjoint passes the dot-product test to 1e-12 (FP64) / 1e-6 (FP32).
"""
import ctypes
import math
from dataclasses import replace

import numpy as np
import pytest
import torch

from gpu_amd import core
from gpu_amd.geometry import Geometry, adjoint_64, multisource_256, plain_sart_128, repo_scale_512
from gpu_amd.phantom import phantom
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


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


GEOMS = {
    "small": small_geom(),
    "offsets": small_geom(x_offset=3, det_offset=2.2, ncols=37, nrows=21),
    "aniso": small_geom(nx=17, ny=23, nz=11, vx=0.9, vy=1.2, vz=0.7, du=0.8, dv=1.3),
    "adjoint_64": adjoint_64().with_views(15),
    "multisource": small_geom(nsources=4, src_z0=-9.0, src_dz=6.0, nrows=12),
    # detector pixels much finer than the voxel footprint: the fast
    # backprojection leaves the LDS-staged kernel for the per-voxel one
    "fine": small_geom(du=0.15, dv=0.15, ncols=200, nrows=150, nviews=5),
    # CT_style 1 and 2 (multi-slice fan beam, parallel beam), rows at the
    # slice centres as the reference sets them up (Z_prj = ZETA, vz = du)
    "fan": small_geom(style=1, nrows=16, dv=1.0),
    "parallel": small_geom(style=2, nrows=16, dv=1.0, det_offset=0.7, x_offset=1),
}


@pytest.mark.parametrize("name", list(GEOMS))
def test_forward_fp32(ctx, name):
    g = GEOMS[name]
    x = phantom(g, seed=11, scale=0.3).numpy()
    ref = orc.forward(g, x, prec=32)
    got = host(ctx.forward(g, dev(x)))
    r = rel(got, ref)
    print(f"{name}: forward FP32 rel L2 {r:.3e}, max abs {np.abs(got - ref).max():.3e}")
    assert r <= 1e-5


@pytest.mark.parametrize("name", list(GEOMS))
def test_exact_mode_is_bit_identical(ctx, name):
    """CT_MATH_EXACT reproduces the reference's FP32 operation sequence: the
    forward projection and two SART sweeps equal the CPU restatement bit for bit."""
    g = GEOMS[name]
    x = phantom(g, seed=21, scale=0.3).numpy()
    ctx.set_math("exact")
    try:
        got = host(ctx.forward(g, dev(x)))
        b = orc.forward(g, x, prec=32)
        v = torch.zeros(g.vol_shape, device="cuda")
        bd = dev(b)
        ref = np.zeros_like(x)
        for _ in range(2):
            ctx.sart_sweep(g, v, bd, beta=0.9, subset=2, order="interleaved")
            ref = orc.sart_sweep(g, ref, b, beta=0.9, subset=2, order="interleaved", prec=32)
        rec = host(v)
    finally:
        ctx.set_math("fast")
    assert np.array_equal(got, b.astype(np.float64))
    assert np.array_equal(rec, ref.astype(np.float64))


def test_fast_forward_accuracy_vs_fp64(ctx):
    """Fast mode (FMA-form plane crossings, FP64 ray sums) at 128^3 / 256^2:
    error against the FP64 oracle within 1.25x of the plain FP32
    restatement's and under 2e-6, five times inside the 1e-5 projection bar.
    The residual ~1e-6 comes from FP32 crossing positions (alpha spans the
    whole source-detector line), which the FP64 sums do not change."""
    g = plain_sart_128().with_views(4)
    x = phantom(g, seed=0).numpy()
    p64 = orc.forward(g, x.astype(np.float64), prec=64)
    p32 = orc.forward(g, x, prec=32)
    got = host(ctx.forward(g, dev(x)))
    e_gpu, e_plain = rel(got, p64), rel(p32, p64)
    print(f"fast vs FP64 {e_gpu:.3e}, plain FP32 vs FP64 {e_plain:.3e}")
    assert e_gpu <= 1.25 * e_plain and e_gpu <= 2e-6


@pytest.mark.parametrize("name", ["small", "aniso", "multisource", "fan", "parallel"])
def test_forward_fp64(ctx, name):
    g = GEOMS[name]
    x = phantom(g, seed=12, scale=0.3).numpy().astype(np.float64)
    ref = orc.forward(g, x, prec=64)
    got = host(ctx.forward(g, dev(x, torch.float64)))
    assert rel(got, ref) <= 1e-12


def test_forward_plain_sart_geometry_view_subset(ctx):
    g = plain_sart_128()
    x = phantom(g, seed=0).numpy()
    for v0 in (0, 44, 123):
        ref = orc.forward(g, x, prec=32, view0=v0, nviews=3)
        got = host(ctx.forward(g, dev(x), view0=v0, nviews=3))
        r = rel(got, ref)
        print(f"view0 {v0}: rel {r:.3e}")
        assert r <= 1e-5


@pytest.mark.parametrize("name", ["small", "offsets", "multisource", "fan", "parallel"])
def test_sart_correction(ctx, name):
    g = GEOMS[name]
    x = phantom(g, seed=13, scale=0.3).numpy()
    meas = orc.forward(g, x, prec=32, view0=3, nviews=2) * 1.1 + 0.01
    ref = orc.sart_correction(g, x * 0.9, meas, view0=3, nviews=2, prec=32)
    got = host(ctx.sart_correction(g, dev(x * 0.9), dev(meas), view0=3, nviews=2))
    assert rel(got, ref) <= 1e-5


@pytest.mark.parametrize("name", list(GEOMS))
def test_sart_correction_batch(ctx, name):
    """Five views: the batched path (z-quad volume, one workgroup per column
    walking the column's shared x/y crossing list).  b = 2 Ax + 0.01 keeps
    b - 0.9 Ax away from cancellation, so the bar measures the ray sums and
    not the FP32 oracle's own rounding amplified by b - Ax."""
    g = GEOMS[name]
    x = phantom(g, seed=16, scale=0.3).numpy()
    meas = orc.forward(g, x, prec=32, view0=0, nviews=5) * 2.0 + 0.01
    ref = orc.sart_correction(g, x * 0.9, meas, view0=0, nviews=5, prec=32)
    got = host(ctx.sart_correction(g, dev(x * 0.9), dev(meas), view0=0, nviews=5))
    r = rel(got, ref)
    print(f"{name}: batched correction rel L2 {r:.3e}")
    assert r <= 1e-5


@pytest.mark.parametrize("angle0", [0.0, 90.0])
def test_forward_full_size(ctx, angle0):
    """The benchmark geometry itself (512^3 at 0.5 mm, 1024^2 at 0.4 mm): four
    views from 0 degrees (rays along x: the column walk, four 256-row
    workgroups per column) and from 90 degrees (the per-ray walk).

    At this size the reference's FP32 setup starts a few rays that do cross
    the volume outside the grid (exact ties at the first view of each
    orbit: 1 ray at 0 degrees, 19 at 90), and its walk returns 0 for them.
    Both modes reproduce that: exact mode bit for bit, fast mode within the
    1e-5 bar over every ray.  Where the reference returns 0, fast mode stays
    below 1e-5 (0 for the tie rays, whose line integrals are 0.04-1.56; at
    most a value times a sub-threshold chord for grazing rays)."""
    g = replace(repo_scale_512(), angle0_deg=angle0).with_views(4)
    x = phantom(g, seed=0, scale=2.5, device="cuda").cpu().numpy()
    ref = orc.forward(g, x, prec=32)
    xd = dev(x)
    got = host(ctx.forward(g, xd))
    ctx.set_math("exact")
    try:
        ex = host(ctx.forward(g, xd))
    finally:
        ctx.set_math("fast")
    assert np.array_equal(ex, ref.astype(np.float64))
    r = rel(got, ref)
    zeroed = ref == 0
    print(f"512^3 from {angle0} deg: fast rel L2 {r:.3e} over all {ref.size} rays; "
          f"{int(zeroed.sum())} rays zero in the reference, max fast value there {got[zeroed].max():.3e}")
    assert r <= 1e-5
    assert got[zeroed].max() <= 1e-5


def test_sart_sweep_full_size(ctx):
    """Four view-by-view SART updates at the benchmark geometry (views 0-1
    walk by column, views from 90 degrees ray by ray; the z-quad volume and
    the LDS-staged update at their real shapes), from the phantom's
    projections and a zero volume, against the oracle at the 1e-4
    reconstruction bar."""
    g0 = repo_scale_512()
    x = phantom(g0, seed=0, scale=2.5, device="cuda").cpu().numpy()
    for angle0 in (0.0, 90.0):
        g = replace(g0, angle0_deg=angle0).with_views(2)
        b = orc.forward(g, x, prec=32)
        ref = orc.sart_sweep(g, np.zeros_like(x), b, beta=0.0625, subset=1, prec=32)
        v = torch.zeros(g.vol_shape, device="cuda")
        ctx.sart_sweep(g, v, dev(b), beta=0.0625, subset=1)
        r = rel(host(v), ref)
        print(f"512^3 sweep from {angle0} deg: rel L2 {r:.3e}")
        assert r <= 1e-4


@pytest.mark.parametrize("math_mode", ["exact", "fast"])
@pytest.mark.parametrize("name", ["small", "offsets", "aniso", "multisource", "fine", "fan", "parallel"])
@pytest.mark.parametrize("mode,nv", [(1, 1), (0, 1), (1, 3)])
def test_backproject_voxel(ctx, name, mode, nv, math_mode):
    """exact: the reference's FP32 operation sequence, bit-identical to the
    oracle; fast: hardware reciprocal/sqrt, held to the 1e-5 bar."""
    g = GEOMS[name]
    rng = np.random.default_rng(5)
    x = phantom(g, seed=14, scale=0.3).numpy()
    corr = rng.standard_normal((nv, *g.proj_shape[1:])).astype(np.float32) * 0.01
    ref = orc.backproject_voxel(g, x, corr, view0=2, nviews=nv, beta=0.7, mode=mode, prec=32)
    v = dev(x)
    ctx.set_math(math_mode)
    try:
        ctx.backproject_voxel(g, v, dev(corr), view0=2, nviews=nv, beta=0.7, mode=mode)
        got = host(v)
    finally:
        ctx.set_math("fast")
    r = rel(got, ref)
    print(f"{name} mode {mode} nv {nv} {math_mode}: rel {r:.3e} max {np.abs(got - ref).max():.3e}")
    if math_mode == "exact":
        assert r == 0.0
    else:
        assert r <= 1e-5


def test_backproject_voxel_fp64(ctx):
    g = GEOMS["aniso"]
    rng = np.random.default_rng(6)
    x = rng.standard_normal(g.vol_shape)
    corr = rng.standard_normal((3, *g.proj_shape[1:]))
    ref = orc.backproject_voxel(g, x, corr, view0=1, nviews=3, beta=0.3, mode=1, prec=64)
    v = dev(x, torch.float64)
    ctx.backproject_voxel(g, v, dev(corr, torch.float64), view0=1, nviews=3, beta=0.3, mode=1)
    assert rel(host(v), ref) <= 1e-12


@pytest.mark.parametrize("name,subset,order", [("small", 1, "ascending"), ("offsets", 3, "interleaved"),
                                               ("multisource", 1, "ascending"), ("aniso", 4, "ascending"),
                                               ("fan", 1, "ascending"), ("parallel", 2, "interleaved")])
def test_sart_sweep(ctx, name, subset, order):
    g = GEOMS[name]
    x = phantom(g, seed=15, scale=0.3).numpy()
    b = orc.forward(g, x, prec=32)
    ref = np.zeros_like(x)
    v = torch.zeros(g.vol_shape, device="cuda")
    bd = dev(b)
    for _ in range(2):
        ref = orc.sart_sweep(g, ref, b, beta=0.9, subset=subset, order=order, prec=32)
        ctx.sart_sweep(g, v, bd, beta=0.9, subset=subset, order=order)
    r = rel(host(v), ref)
    print(f"{name}: SART rel {r:.3e}")
    assert r <= 1e-4


def test_sart_plain_128_two_iterations(ctx):
    """128^3 / 180 views / 256^2, relaxation 1.0, subsets of 20 views (interleaved),
    2 iterations against the FP32 oracle; error vs the FP64 oracle reported."""
    g = plain_sart_128()
    x = phantom(g, seed=0).numpy()
    b = orc.forward(g, x, prec=32)
    v = torch.zeros(g.vol_shape, device="cuda")
    bd = dev(b)
    ref = np.zeros_like(x)
    for _ in range(2):
        ctx.sart_sweep(g, v, bd, beta=1.0, subset=20, order="interleaved")
        ref = orc.sart_sweep(g, ref, b, beta=1.0, subset=20, order="interleaved", prec=32)
    got = host(v)
    r = rel(got, ref)
    print(f"plain_sart_128: GPU vs FP32 oracle rel {r:.3e}; vs truth rel {rel(got, x):.3e}")
    assert r <= 1e-4


@pytest.mark.parametrize("name", ["small", "multisource", "fan", "parallel"])
def test_adjoint_matches_oracle(ctx, name):
    g = GEOMS[name]
    rng = np.random.default_rng(8)
    y = rng.standard_normal(g.proj_shape)
    ref64 = orc.adjoint(g, y, prec=64)
    got64 = host(ctx.adjoint(g, dev(y, torch.float64)))
    assert rel(got64, ref64) <= 1e-12
    ref32 = orc.adjoint(g, y, prec=32)
    got32 = host(ctx.adjoint(g, dev(y)))
    assert rel(got32, ref32) <= 1e-5


def test_adjoint_multisource_full_size(ctx):
    """The multi-source autograd geometry itself (48 sources 3 mm apart, 1024 x
    128 detector, 256^3 at 0.75 mm): the sources' spread needs 36 footprint
    rows per 16-voxel tile, which the gather adjoint's 48-row window holds.
    One view (48 images) of N(0, 1) data against the FP64 oracle.  A voxel's
    segment is the difference of two FP32 plane crossings of size ~0.5 (6e-8
    rounding) over an alpha length of ~5e-4 (0.75 mm of a 1.5 m ray), so each
    segment carries ~1.2e-4 relative error.  The forward's sums telescope it
    away; a voxel's adjoint sum of unrelated rays of cancelling N(0, 1) data
    does not.  Hence the 2e-4 bar, against the FP64 oracle.  The matched
    pair itself passes the FP32 dot test over two views."""
    g = multisource_256().with_views(1)
    rng = np.random.default_rng(9)
    y = rng.standard_normal(g.proj_shape).astype(np.float32)
    ref64 = orc.adjoint(g, y.astype(np.float64), prec=64)
    ref32 = orc.adjoint(g, y, prec=32)
    got32 = host(ctx.adjoint(g, dev(y)))
    e_gpu, e_plain = rel(got32, ref64), rel(ref32, ref64)
    lhs, rhs = ctx.adjoint_test(multisource_256().with_views(2), dtype=torch.float32, seed=5)
    d = abs(lhs - rhs) / abs(lhs)
    print(f"multi-source 256^3: adjoint vs FP64 {e_gpu:.3e} (plain FP32 {e_plain:.3e}, "
          f"GPU vs plain {rel(got32, ref32):.3e}), dot test {d:.3e}")
    assert e_gpu <= 2e-4
    assert d <= 1e-6


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-6)])
def test_adjoint_dot_product_adjoint_64(ctx, dtype, tol):
    g = adjoint_64()
    lhs, rhs = ctx.adjoint_test(g, dtype=dtype, seed=2024)
    err = abs(lhs - rhs) / abs(lhs)
    print(f"{dtype}: <Ax,y>={lhs:.12e} <x,A^T y>={rhs:.12e} rel {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("name", ["small", "offsets", "multisource", "fine", "fan", "parallel"])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-6)])
def test_adjoint_dot_product_paths(ctx, name, dtype, tol):
    """Fast-mode FP32 adjoints: the gather form (every voxel sums the walk's
    own segments over its detector footprint, no atomics) on the small
    geometries, the atomic scatter of the same walk on "fine" (footprints too
    wide for the gather's LDS window); FP64 takes the reference traversal."""
    g = GEOMS[name]
    lhs, rhs = ctx.adjoint_test(g, dtype=dtype, seed=77)
    err = abs(lhs - rhs) / abs(lhs)
    print(f"{name} {dtype}: rel {err:.3e}")
    assert err <= tol


def test_power_iteration_vs_dense_svd(ctx):
    g = Geometry(nx=6, ny=6, nz=4, vx=1.0, ncols=9, nrows=6, du=1.2, nviews=5, dso=40.0, dod=-20.0)
    n = g.nvox
    A = np.zeros((g.rays, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        A[:, i] = orc.forward(g, e.reshape(g.vol_shape), prec=64).ravel()
    smax2 = np.linalg.svd(A, compute_uv=False)[0] ** 2
    lip = ctx.power_iteration(g, dtype=torch.float64, niter=300, seed=1)
    assert lip == pytest.approx(smax2, rel=1e-6)


def test_fista_momentum(ctx):
    rng = np.random.default_rng(1)
    x, xp = rng.standard_normal((8, 9, 10)), rng.standard_normal((8, 9, 10))
    tn, gm = core.Context.fista_gamma(2.5)
    assert tn == pytest.approx((1 + math.sqrt(1 + 4 * 6.25)) / 2)
    for dt, prec in ((torch.float32, 32), (torch.float64, 64)):
        X, P = dev(x, dt), dev(xp, dt)
        ctx.fista_momentum(X, P, X, gm)
        y_ref, p_ref = orc.fista_momentum(x, xp, gm, prec=prec)
        assert rel(host(X), y_ref) <= (1e-7 if prec == 32 else 1e-15)
        assert rel(host(P), p_ref) == 0.0


@pytest.mark.parametrize("shape", [(7, 9, 11), (16, 24, 32), (131, 19, 70)])
def test_tv_fgp(ctx, shape):
    """(131, 19, 70): several z chunks and partial x/y tiles of the fused step."""
    rng = np.random.default_rng(4)
    b = (rng.standard_normal(shape) * 0.01 + 0.02).astype(np.float32)
    ref = orc.tv_fgp(b, 0.004, 20, prec=32)
    v = dev(b)
    ctx.tv_fgp(v, 0.004, 20)
    r = rel(host(v), ref)
    print(f"FGP {shape}: rel {r:.3e}")
    assert r <= 1e-5
    ref64 = orc.tv_fgp(b.astype(np.float64), 0.004, 20, prec=64)
    v = dev(b, torch.float64)
    ctx.tv_fgp(v, 0.004, 20)
    assert rel(host(v), ref64) <= 1e-12


def test_tv_difference_pair(ctx):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((6, 7, 8))
    r, s, t = ctx.tv_fgp_D(dev(x, torch.float64))
    rr, ss, tt = orc.tv_D(x)
    for a, b in ((r, rr), (s, ss), (t, tt)):
        assert rel(host(a), b) <= 1e-15
    p = [rng.standard_normal((6, 7, 8)) for _ in range(3)]
    out = host(ctx.tv_fgp_Dt(*(dev(q, torch.float64) for q in p)))
    assert rel(out, orc.tv_Dt(*p)) <= 1e-15


def test_norms(ctx):
    rng = np.random.default_rng(6)
    a, b = rng.standard_normal((9, 10, 11)), rng.standard_normal((9, 10, 11))
    for dt, prec, tol in ((torch.float64, 64, 1e-13), (torch.float32, 32, 1e-12)):
        A, B = dev(a, dt), dev(b, dt)
        assert ctx.residual_norm2(A, B) == pytest.approx(orc.residual_norm2(host(A), host(B), prec=64), rel=tol)
        assert ctx.tv_norm(A) == pytest.approx(orc.tv_norm(host(A), prec=64), rel=tol)
        assert ctx.tv_eps(A, 1e-8) == pytest.approx(orc.tv_eps(host(A), 1e-8), rel=tol)
        assert ctx.dot(A, B) == pytest.approx(float((host(A) * host(B)).sum()), rel=1e-11)


@pytest.mark.parametrize("name", ["small", "multisource", "parallel"])
def test_objective(ctx, name):
    """Fused residual + TV norm equals the oracle's forward projection followed
    by the FP64 norms (V63/main_recon.cu:401-428)."""
    g = GEOMS[name]
    x = phantom(g, seed=19, scale=0.3).numpy()
    b = (orc.forward(g, x, prec=64) * 1.03 + 0.01).astype(np.float32)
    y = (x * 0.9).astype(np.float32)
    for dt, prec, tol in ((torch.float32, 32, 1e-5), (torch.float64, 64, 1e-12)):
        data, tv, obj = ctx.objective(g, dev(y, dt), dev(b, dt), lambda_tv=0.05)
        ax = orc.forward(g, y.astype(orc._np_dtype(prec)), prec=prec)
        ref_data = orc.residual_norm2(ax, b.astype(orc._np_dtype(prec)), prec=64)
        ref_tv = orc.tv_norm(y.astype(np.float64), prec=64)
        assert data == pytest.approx(ref_data, rel=tol)
        assert tv == pytest.approx(ref_tv, rel=1e-12)
        assert obj == pytest.approx(data + 0.1 * tv, rel=1e-15)


def test_tv_eps_grad(ctx):
    x = np.random.default_rng(7).standard_normal((5, 6, 7)) * 0.1
    got = host(ctx.tv_eps_grad(dev(x, torch.float64), 1e-8))
    assert rel(got, orc.tv_eps_grad(x, 1e-8)) <= 1e-12


FDK_G = Geometry(nx=48, ny=48, nz=24, vx=1.0, ncols=96, nrows=40, du=1.0, nviews=60, dso=300.0, dod=-150.0)


@pytest.mark.parametrize("window", ["ramp", "hann"])
def test_fdk_stages(ctx, window):
    g = FDK_G
    x = phantom(g, seed=3, scale=0.4).numpy()
    p = orc.forward(g, x, prec=32)
    w_ref = orc.fdk_weight(g, p, prec=32)
    w = ctx.fdk_weight(g, dev(p))
    assert rel(host(w), w_ref) <= 1e-6
    f_ref = orc.ramp_filter(g, w_ref, window, prec=32)
    f = ctx.ramp_filter(g, dev(w_ref), window=window)
    r = rel(host(f), f_ref)
    print(f"ramp {window}: rel {r:.3e}")
    assert r <= 1e-5
    v_ref = orc.fdk_backproject(g, f_ref, np.zeros(g.vol_shape, np.float32), prec=32)
    v = ctx.fdk_backproject(g, dev(f_ref), torch.zeros(g.vol_shape, device="cuda"))
    r = rel(host(v), v_ref)
    print(f"fdk backprojection: rel {r:.3e}")
    assert r <= 1e-5
    full_ref = orc.fdk(g, p, window, prec=32)
    full = ctx.fdk(g, dev(p), window=window)
    r = rel(host(full), full_ref)
    print(f"fdk {window}: rel {r:.3e}")
    assert r <= 1e-4


FDK_PATHS = {
    # 16x16-column tiles whose detector windows fit the LDS staging; nz = 24
    # leaves a part tile of 8 voxels along z
    "staged": FDK_G,
    # detector pixels 3x finer than the voxels: tile windows wider than 64
    # pixels, gathered from global memory
    "wide": Geometry(nx=40, ny=40, nz=20, vx=1.0, ncols=256, nrows=120, du=0.3, nviews=24, dso=300.0, dod=-150.0),
    # off-centre grid and detector, grid sides that are not tile multiples
    "offsets": Geometry(nx=40, ny=36, nz=20, vx=1.0, ncols=96, nrows=40, du=1.0, nviews=24, dso=300.0, dod=-150.0,
                        x_offset=3, det_offset=2.5),
    # a short scan (24 x 5.5 degrees): the view weight is not 1 per degree
    "short_scan": Geometry(nx=40, ny=40, nz=20, vx=1.0, ncols=96, nrows=40, du=1.0, nviews=24, dso=300.0,
                           dod=-150.0, angle_step_deg=5.5, angle0_deg=12.0),
}


@pytest.mark.parametrize("name", list(FDK_PATHS))
def test_fdk_backprojection_paths(ctx, name):
    """The FDK backprojector on realistic filtered views, over a view range
    with an odd count (a part batch) into a non-zero volume (FDK accumulates):
    fast mode (LDS-staged windows, global-memory gathers for wide windows and
    FP16 views) within 1e-5 of the FP32 restatement, exact mode bit-identical."""
    g = FDK_PATHS[name]
    x = phantom(g, seed=4, scale=0.4).numpy()
    p = orc.forward(g, x, prec=32)
    f = orc.ramp_filter(g, orc.fdk_weight(g, p, prec=32), "hann", prec=32)
    v0 = phantom(g, seed=6, scale=0.2).numpy()
    view0, nv = 3, (g.nviews - 6) | 1
    ref = orc.fdk_backproject(g, f, v0, view0=view0, nviews=nv, prec=32)
    fast = host(ctx.fdk_backproject(g, dev(f), dev(v0), view0=view0, nviews=nv))
    r = rel(fast, ref)
    print(f"{name}: fdk backprojection fast rel {r:.3e}")
    assert r <= 1e-5
    f16 = f.astype(np.float16)
    ref16 = orc.fdk_backproject(g, f16.astype(np.float32), v0, view0=view0, nviews=nv, prec=32)
    got16 = host(ctx.fdk_backproject(g, dev(f16, torch.float16), dev(v0), view0=view0, nviews=nv))
    r16 = rel(got16, ref16)
    print(f"{name}: FP16 views rel {r16:.3e}")
    assert r16 <= 1e-5
    ctx.set_math("exact")
    try:
        exact = host(ctx.fdk_backproject(g, dev(f), dev(v0), view0=view0, nviews=nv))
    finally:
        ctx.set_math("fast")
    assert np.array_equal(exact, ref.astype(np.float64))


MS_FDK_G = Geometry(nx=32, ny=30, nz=24, vx=1.0, ncols=64, nrows=24, du=1.0, nviews=12, dso=200.0, dod=-100.0,
                    nsources=3, src_z0=-8.0, src_dz=8.0)


@pytest.mark.parametrize("sources", ["inverse_square", "nearest", "equal"])
def test_multisource_fdk(ctx, sources):
    """Multi-source FDK (TB/FDK): the backprojection with each source
    combination is bit-identical to the FP32 restatement in both arithmetic
    modes (one kernel, the reference's operation order) and matches FP64 to
    1e-12; the whole pipeline (per-source cosine weight, filter of every
    image, combination) is within 1e-4 of the restatement."""
    g = MS_FDK_G
    rng = np.random.default_rng(11)
    f = rng.standard_normal(g.proj_shape).astype(np.float32)
    v0 = phantom(g, seed=3, scale=0.2).numpy()
    ref = orc.fdk_backproject_ms(g, f, v0, mode=sources, view0=2, nviews=7, prec=32)
    for mode in ("fast", "exact"):
        ctx.set_math(mode)
        try:
            got = host(ctx.fdk_backproject(g, dev(f), dev(v0), view0=2, nviews=7, sources=sources))
        finally:
            ctx.set_math("fast")
        assert np.array_equal(got, ref.astype(np.float64)), mode
    ref64 = orc.fdk_backproject_ms(g, f.astype(np.float64), v0.astype(np.float64), mode=sources, view0=2, nviews=7,
                                   prec=64)
    got64 = host(ctx.fdk_backproject(g, dev(f, torch.float64), dev(v0, torch.float64), view0=2, nviews=7,
                                     sources=sources))
    assert rel(got64, ref64) <= 1e-12
    x = phantom(g, seed=4, scale=0.4).numpy()
    p = orc.forward(g, x, prec=32)
    w_ref = orc.fdk_weight(g, p, prec=32)
    assert rel(host(ctx.fdk_weight(g, dev(p))), w_ref) <= 1e-6
    full_ref = orc.fdk(g, p, "hann", prec=32, sources=sources)
    full = host(ctx.fdk(g, dev(p), window="hann", sources=sources))
    r = rel(full, full_ref)
    print(f"multi-source FDK {sources}: rel {r:.3e}")
    assert r <= 1e-4


def test_fdk_is_cone_beam_only(ctx):
    for style in (1, 2):
        g = replace(FDK_G, style=style)
        with pytest.raises(core.CtError) as e:
            ctx.fdk(g, torch.zeros(g.proj_shape, device="cuda"))
        assert e.value.status == 2


def test_ramp_filter_non_power_of_two(ctx):
    g = Geometry(nx=8, ny=8, nz=8, vx=1.0, ncols=45, nrows=7, du=0.9, nviews=3, dso=100.0, dod=-50.0)
    p = np.random.default_rng(2).standard_normal((3, 7, 45)).astype(np.float32)
    ref = orc.ramp_filter(g, p, "hann", prec=64)
    got = host(ctx.ramp_filter(g, dev(p), window="hann"))
    assert rel(got, ref) <= 1e-5


@pytest.mark.parametrize("ncols", [8, 256, 512, 1024, 2048])
def test_ramp_filter_power_of_two_pads(ctx, ncols):
    """Pads 2^9..2^12 take the Stockham radix-8 kernel, 16 the radix-2 LDS
    kernel; an odd number of rows leaves the last pair with one row.  Against
    the FP64 restatement (numpy-pinned) to 1e-5, both windows."""
    g = Geometry(nx=8, ny=8, nz=8, vx=1.0, ncols=ncols, nrows=3, du=0.9, nviews=3, dso=100.0, dod=-50.0)
    p = np.random.default_rng(ncols).standard_normal((3, 3, ncols)).astype(np.float32)
    for window in ("ramp", "hann"):
        ref = orc.ramp_filter(g, p, window, prec=64)
        got = host(ctx.ramp_filter(g, dev(p), window=window))
        r = rel(got, ref)
        print(f"ncols {ncols} {window}: rel {r:.3e}")
        assert r <= 1e-5


def test_fdk_fused_weight_and_filter(ctx):
    """ct_fdk applies the cosine weight while the Stockham filter reads the
    rows (pad 512); the result matches the FP32 restatement of weight, filter
    and backprojection."""
    g = Geometry(nx=32, ny=32, nz=16, vx=1.0, ncols=256, nrows=24, du=0.5, nviews=20, dso=300.0, dod=-150.0)
    x = phantom(g, seed=9, scale=0.4).numpy()
    p = orc.forward(g, x, prec=32)
    ref = orc.fdk(g, p, "hann", prec=32)
    got = host(ctx.fdk(g, dev(p), window="hann"))
    r = rel(got, ref)
    print(f"fused FDK: rel {r:.3e}")
    assert r <= 1e-4


def test_sart_fista_iteration_matches_composition(ctx):
    g = GEOMS["small"]
    x = phantom(g, seed=16, scale=0.3).numpy()
    b = orc.forward(g, x, prec=32)
    X = torch.zeros(g.vol_shape, device="cuda")
    XP = torch.zeros_like(X)
    y_ref, xp_ref, t = np.zeros_like(x), np.zeros_like(x), 1.0
    bd = dev(b)
    t_gpu = 1.0
    for _ in range(3):
        t_gpu = ctx.sart_fista_iteration(g, X, XP, bd, t_gpu, beta=0.8, subset=3, order="interleaved",
                                         fgp_lambda=0.002, fgp_niter=10)
        xk = orc.sart_sweep(g, y_ref, b, beta=0.8, subset=3, order="interleaved", prec=32)
        xk = orc.tv_fgp(xk, 0.002, 10, prec=32)
        tn = (1 + math.sqrt(1 + 4 * t * t)) / 2
        y_ref, xp_ref = orc.fista_momentum(xk, xp_ref, (t - 1) / tn, prec=32)
        t = tn
    assert t_gpu == pytest.approx(t)
    assert rel(host(XP), xp_ref) <= 1e-4
    assert rel(host(X), y_ref) <= 1e-4


@pytest.mark.parametrize("dt", [torch.float16, torch.bfloat16])
def test_reduced_precision_storage(ctx, dt):
    g = GEOMS["small"]
    x = phantom(g, seed=17, scale=0.3).numpy()
    xs = torch.as_tensor(x).to(dt)
    x_rounded = xs.float().numpy()
    ref = orc.forward(g, x_rounded, prec=32)
    got = host(ctx.forward(g, xs.cuda(), out_dtype=torch.float32))
    assert rel(got, ref) <= 1e-6           # FP32 accumulation of the stored values
    got16 = host(ctx.forward(g, xs.cuda()))
    assert rel(got16, ref) <= (1e-3 if dt == torch.float16 else 8e-3)


@pytest.mark.parametrize("dt,tol", [(torch.float16, 3e-3), (torch.bfloat16, 1.2e-2)])
@pytest.mark.parametrize("name", ["small", "offsets", "aniso", "multisource", "fan", "fine"])
@pytest.mark.parametrize("mode,nv", [(0, 1), (1, 1), (1, 3)])
def test_backproject_reduced_precision(ctx, dt, tol, name, mode, nv):
    """FP16/BF16 volumes take the packed-half chord engine: chord intervals in
    voxel-local FP16 units, products accumulated in FP32, and the updated
    value stored with stochastic rounding.  Against the FP32 oracle on the
    storage-rounded input, the error is that of the stored result (up to one
    ulp: FP16 1e-3, BF16 8e-3 relative) plus the engine's ~1e-3 on W/L.
    "fine" takes the per-voxel kernel (footprints too wide for the LDS
    window), whose 16-bit updates are stochastically rounded too."""
    g = GEOMS[name]
    rng = np.random.default_rng(8)
    x = torch.as_tensor(phantom(g, seed=18, scale=0.3).numpy()).to(dt)
    x_rounded = x.float().numpy()
    corr = rng.standard_normal((nv, *g.proj_shape[1:])).astype(np.float32) * 0.01
    ref = orc.backproject_voxel(g, x_rounded, corr, view0=2, nviews=nv, beta=0.7, mode=mode, prec=32)
    v = x.cuda()
    ctx.backproject_voxel(g, v, dev(corr), view0=2, nviews=nv, beta=0.7, mode=mode)
    got = host(v)
    r = rel(got, ref)
    print(f"{name} {dt} mode {mode} nv {nv}: rel {r:.3e}")
    assert r <= tol


def test_backproject_fp16_out_of_range_correction(ctx):
    """One correction pixel far beyond the FP16 range (as a ray that grazes
    the volume, with a tiny length to divide by, can produce): the
    packed-half engine's table clamps finite values to the FP16 range, so the
    16-bit update stays finite everywhere, and voxels whose footprint misses
    that pixel match the FP32 update as in the plain case."""
    g = GEOMS["small"]
    x = torch.as_tensor(phantom(g, seed=18, scale=0.3).numpy()).half()
    corr = np.random.default_rng(9).standard_normal((1, *g.proj_shape[1:])).astype(np.float32) * 0.01
    big = corr.copy()
    big[0, g.nrows // 2, g.ncols // 2] = 1e9
    v16 = x.cuda()
    ctx.backproject_voxel(g, v16, dev(big), view0=2, nviews=1, beta=0.7, mode=1)
    got = host(v16)
    assert np.isfinite(got).all()
    x32 = x.float().numpy()
    ref = orc.backproject_voxel(g, x32, corr, view0=2, nviews=1, beta=0.7, mode=1, prec=32)
    ref_big = orc.backproject_voxel(g, x32, big, view0=2, nviews=1, beta=0.7, mode=1, prec=32)
    away = ref_big == ref           # voxels the large pixel does not reach
    assert away.sum() > 0.5 * away.size and (~away).sum() > 0
    assert rel(got[away], ref[away]) <= 3e-3


@pytest.mark.parametrize("dt,tol", [(torch.float32, 1e-5), (torch.float16, 3e-3)])
@pytest.mark.parametrize("nz,tie", [(16, True), (15, False)])
def test_backproject_horizontal_row(ctx, dt, tol, nz, tie):
    """A detector row level with the source (21 rows, source at z = 0) casts
    horizontal rays, whose 1/ez is infinite.  With 16 slices those rays lie
    exactly on the face between slices 7 and 8, where the reference's min/max
    of (+-inf, NaN) gives neither voxel a chord: a correction on that row
    alone changes nothing (the oracle agrees).  With 15 slices they run
    through the middle of slice 7, whose voxels take full chords.  The
    LDS-staged kernels (FP32 and the FP16 packed-half engine) resolve both
    from their ordered row table: at the tie the volume comes back
    bit-identical (a chord given to either voxel would move it by ~0.007),
    otherwise only slice 7 moves and the volume matches the oracle's to the
    storage type's bar."""
    g = small_geom(nz=nz, nrows=21)
    xs = torch.as_tensor(phantom(g, seed=14, scale=0.3).numpy()).to(dt)
    x_in = xs.float().numpy()
    corr = np.zeros((1, *g.proj_shape[1:]), np.float32)
    corr[0, g.nrows // 2, :] = 0.01
    ref = orc.backproject_voxel(g, x_in, corr, view0=2, nviews=1, beta=0.7, mode=1, prec=32)
    v = xs.cuda()
    ctx.backproject_voxel(g, v, dev(corr), view0=2, nviews=1, beta=0.7, mode=1)
    got = host(v)
    if tie:
        assert np.array_equal(ref, x_in)
        assert np.array_equal(got, x_in)
    else:
        moved = ref != x_in
        assert moved.sum() > 100 and not moved[:7].any() and not moved[8:].any()
        assert np.array_equal(got[:7], x_in[:7]) and np.array_equal(got[8:], x_in[8:])
        assert rel(got, ref) <= tol


@pytest.mark.parametrize("path", ["exact", "per_voxel"])
@pytest.mark.parametrize("nz,tie", [(16, True), (15, False)])
def test_backproject_horizontal_row_other_paths(ctx, path, nz, tie):
    """The horizontal-row case of test_backproject_horizontal_row on the
    other two voxel backprojectors (FP32): exact mode, bit-identical to the
    oracle, and the per-voxel fast kernel, which takes footprints too wide
    for the LDS window (detector pixels of 0.15 voxel; 151 rows, the middle
    one level with the source).  At the tie the volume is unchanged;
    otherwise only slice 7 moves."""
    if path == "exact":
        g = small_geom(nz=nz, nrows=21)
    else:
        g = small_geom(nz=nz, du=0.15, dv=0.15, ncols=200, nrows=151, nviews=5)
    x = phantom(g, seed=14, scale=0.3).numpy()
    corr = np.zeros((1, *g.proj_shape[1:]), np.float32)
    corr[0, g.nrows // 2, :] = 0.01
    ref = orc.backproject_voxel(g, x, corr, view0=2, nviews=1, beta=0.7, mode=1, prec=32)
    v = dev(x)
    ctx.set_math("exact" if path == "exact" else "fast")
    try:
        ctx.backproject_voxel(g, v, dev(corr), view0=2, nviews=1, beta=0.7, mode=1)
        got = host(v)
    finally:
        ctx.set_math("fast")
    if tie:
        assert np.array_equal(ref, x) and np.array_equal(got, x)
    else:
        moved = ref != x
        assert moved.sum() > 100 and not moved[:7].any() and not moved[8:].any()
        assert np.array_equal(got[:7], x[:7]) and np.array_equal(got[8:], x[8:])
        if path == "exact":
            assert np.array_equal(got, ref)
        else:
            assert rel(got, ref) <= 1e-5


@pytest.mark.parametrize("name", ["small", "fine"])
@pytest.mark.parametrize("x0", [3e-6, 0.01])
def test_stochastic_rounding_is_unbiased(ctx, name, x0):
    """64 single-view FP16 updates of 0.39 storage ulp each (constant
    correction, so every hit voxel gains beta*kappa per view): with
    stochastic rounding the mean gain over the hit voxels is the exact one to
    within 5%, in the FP16 subnormal range (x0 = 3e-6, ulp 2^-24) and in the
    normal range (x0 = 0.01, ulp 2^-17); rounding to nearest would keep
    none of it.  The LDS-staged kernel ("small") and the per-voxel kernel
    ("fine") both store this way."""
    g = GEOMS[name]
    ulp = 2.0 ** -24 if x0 < 6.1e-5 else 2.0 ** -17
    kappa = 0.37
    beta = 0.39 * ulp / kappa
    v = torch.full(g.vol_shape, x0, device="cuda", dtype=torch.float16)
    start = host(v)
    corr = torch.full((1, *g.proj_shape[1:]), kappa, device="cuda")
    nup = 64
    for j in range(nup):
        ctx.backproject_voxel(g, v, corr, view0=j % g.nviews, nviews=1, beta=beta, mode=1)
    got = host(v) - start
    # the exact gain: beta*kappa for every view whose update reaches the voxel
    ones = np.ones((1, *g.proj_shape[1:]))
    nhit = sum((orc.backproject_voxel(g, np.zeros(g.vol_shape), ones, view0=j % g.nviews, nviews=1, beta=1.0,
                                      mode=1, prec=64) > 0).astype(np.int64) for j in range(nup))
    hit = nhit > 0
    assert hit.sum() > 100
    want = float((nhit[hit] * beta * kappa).mean())
    mean = float(got[hit].mean())
    print(f"{name} x0 {x0}: mean gain {mean:.4e}, exact {want:.4e} ({mean / want:.4f})")
    assert abs(mean / want - 1) < 0.05


def test_compat_layer(ctx):
    lib = core.lib()
    g = GEOMS["small"]
    gc = g.to_c()
    lib.ct_compat_set_geometry.argtypes = [ctypes.POINTER(type(gc))]
    lib.ct_compat_set_geometry.restype = ctypes.c_int
    assert lib.ct_compat_set_geometry(ctypes.byref(gc)) == 0
    x = phantom(g, seed=18, scale=0.3).numpy().astype(np.float32)
    proj = np.zeros(g.proj_shape, dtype=np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.Forward_3D_ray_driven_siddon.argtypes = [fp, fp]
    lib.Forward_3D_ray_driven_siddon(x.ctypes.data_as(fp), proj.ctypes.data_as(fp))
    assert lib.ct_compat_last_status() == 0
    assert rel(proj, orc.forward(g, x, prec=32)) <= 1e-5
    vol = np.zeros_like(x)
    lib.Reconstruction_3D_ray_driven_CBCT.argtypes = [fp, fp, ctypes.c_float]
    lib.Reconstruction_3D_ray_driven_CBCT(vol.ctypes.data_as(fp), proj.ctypes.data_as(fp), ctypes.c_float(0.5))
    assert rel(vol, orc.sart_sweep(g, np.zeros_like(x), proj, beta=0.5, subset=1, prec=32)) <= 1e-4
    a = proj.copy()
    b = proj * 0.5
    lib.L2_norm_gpu.argtypes = [fp, fp]
    lib.L2_norm_gpu.restype = ctypes.c_double
    v = lib.L2_norm_gpu(a.ctypes.data_as(fp), b.ctypes.data_as(fp))
    assert v == pytest.approx(float(((proj.astype(np.float64) * 0.5) ** 2).sum()), rel=1e-6)
    np.testing.assert_allclose(a, (proj - b) ** 2, rtol=1e-6)
    lib.ct_compat_release()


def test_volume_over_4gib_takes_64bit_path(ctx):
    """A 1032 x 1032 x 1016 FP32 volume (4.3 GB, past the 32-bit buffer
    offsets): the fast forward switches to 64-bit addressing with an explicit
    range test; it must agree with the exact mode (the reference's traversal)
    to the projection bar.  A small detector keeps the oracle out of it."""
    g = Geometry(nx=1032, ny=1032, nz=1016, vx=0.25, ncols=24, nrows=16, du=3.0, nviews=3,
                 dso=1000.0, dod=-500.0, angle_step_deg=31.0)
    assert g.nvox * 4 > (1 << 32)
    x = phantom(g, seed=5, scale=1.0, device="cuda")
    fast = host(ctx.forward(g, x))
    ctx.set_math("exact")
    try:
        exact = host(ctx.forward(g, x))
    finally:
        ctx.set_math("fast")
    del x
    torch.cuda.empty_cache()
    r = rel(fast, exact)
    print(f"4.3 GB volume: fast vs exact rel {r:.3e}")
    assert np.abs(exact).max() > 0 and r <= 1e-5


def test_empty_and_degenerate_inputs(ctx):
    """Zero views are a no-op; a zero projection gives a zero adjoint and
    leaves a SART update unchanged; rays that miss the volume give 0."""
    g = GEOMS["small"]
    x = dev(phantom(g, seed=3, scale=0.3).numpy())
    before = x.clone()
    lib, gc = core.lib(), g.to_c()
    p = torch.full((1, *g.proj_shape[1:]), 7.0, device="cuda")
    assert lib.ct_forward(ctx._h, ctypes.byref(gc), ctypes.c_void_p(x.data_ptr()), 0,
                          ctypes.c_void_p(p.data_ptr()), 0, 3, 0) == 0
    assert lib.ct_backproject_voxel(ctx._h, ctypes.byref(gc), ctypes.c_void_p(x.data_ptr()), 0,
                                    ctypes.c_void_p(p.data_ptr()), 3, 0, ctypes.c_double(1.0), 1) == 0
    torch.cuda.synchronize()
    assert torch.equal(x, before) and bool((p == 7.0).all())
    z = torch.zeros((g.nviews, *g.proj_shape[1:]), device="cuda")
    assert float(ctx.adjoint(g, z).abs().max()) == 0.0
    ctx.backproject_voxel(g, x, z[:2], view0=0, nviews=2, beta=1.0, mode=1)
    assert torch.equal(x, before)
    # a detector far off to the side: every ray misses the volume
    gm = small_geom(det_offset=500.0)
    assert float(ctx.forward(gm, x).abs().max()) == 0.0
