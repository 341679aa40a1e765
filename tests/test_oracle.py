"""Pins the CPU oracle against mathematics that fixes each function.

The reference ships no golden vectors or known-answer tests for this path
(its only check, CheckMatchedJointOperator, prints two inner products), so the
oracle is pinned by independent FP64 computations: a brute-force Siddon that
sorts every plane crossing, analytic ray-box chords, the inverse-crime fixed
point, dot-product identities, finite differences, numpy.fft and closed forms.
"""
import math

import numpy as np
import pytest

from gpu_amd.geometry import Geometry, adjoint_64
from gpu_amd.phantom import phantom
from oracle import oracle as orc


def small_geom(**kw):
    base = dict(nx=12, ny=10, nz=8, vx=1.0, ncols=24, nrows=16, du=1.0, nviews=7, dso=60.0, dod=-30.0,
                angle_step_deg=360.0 / 7 + 0.37)
    base.update(kw)
    return Geometry(**base)


def ray_endpoints(g, view, col, row, src=0):
    s, c = orc.view_angle(g, view, prec=64)
    vy = g.vy or g.vx
    sz = g.src_z0 + src * g.src_dz if g.nsources > 1 else g.src_z0
    u = -g.du * (g.ncols / 2 - 0.5) + g.det_offset + col * g.du
    dv = g.dv or g.du
    v = -dv * (g.nrows / 2 - 0.5) + row * dv
    x1 = np.array([g.dod * c - u * s, g.dod * s + u * c, v])
    if g.style == 0:      # cone beam: one source
        x2 = np.array([g.dso * c - g.det_offset * s, g.dso * s + g.det_offset * c, sz])
    elif g.style == 1:    # multi-slice fan beam: the source sits at the row's height
        x2 = np.array([g.dso * c - g.det_offset * s, g.dso * s + g.det_offset * c, v])
    else:                 # parallel beam: the source moves with the pixel
        x2 = np.array([g.dso * c - u * s, g.dso * s + u * c, v])
    return x1, x2


def brute_force_segments(g, x1, x2):
    """Siddon (1985) by sorting every plane crossing, FP64: {voxel index: length}."""
    vy, vz = g.vy or g.vx, g.vz or g.vx
    b = np.array([-g.vx * (g.nx / 2 + g.x_offset), -vy * g.ny / 2, -vz * g.nz / 2])
    h = np.array([g.vx, vy, vz])
    n = np.array([g.nx, g.ny, g.nz])
    d = x2 - x1
    lo, hi = 0.0, 1.0
    alphas = [0.0, 1.0]
    for a in range(3):
        if abs(d[a]) < 1e-12:
            if not (b[a] <= x1[a] <= b[a] + h[a] * n[a]):
                return {}
            continue
        t0 = (b[a] - x1[a]) / d[a]
        t1 = (b[a] + h[a] * n[a] - x1[a]) / d[a]
        lo, hi = max(lo, min(t0, t1)), min(hi, max(t0, t1))
        planes = b[a] + h[a] * np.arange(n[a] + 1)
        alphas.extend(((planes - x1[a]) / d[a]).tolist())
    if hi <= lo:
        return {}
    al = np.array(sorted(set(x for x in alphas if lo <= x <= hi)))
    L = np.linalg.norm(d)
    out = {}
    for a0, a1 in zip(al[:-1], al[1:]):
        if a1 - a0 <= 1e-14:
            continue
        mid = x1 + 0.5 * (a0 + a1) * d
        idx = np.floor((mid - b) / h).astype(int)
        if np.any(idx < 0) or np.any(idx >= n):
            continue
        lin = (idx[2] * g.ny + idx[1]) * g.nx + idx[0]
        out[lin] = out.get(lin, 0.0) + L * (a1 - a0)
    return out


def oracle_segments(g, view, col, row, src=0):
    idx, ln = orc.ray_segments(g, view, col, row, src, prec=64)
    out = {}
    for i, l in zip(idx.tolist(), ln.tolist()):
        out[i] = out.get(i, 0.0) + l
    return {k: v for k, v in out.items() if abs(v) > 1e-12}


@pytest.mark.parametrize("geom_kw", [
    {},
    dict(x_offset=2, det_offset=1.3),
    dict(nx=9, ny=13, nz=5, vx=0.7, vy=1.1, vz=0.9, du=0.8, dv=1.2),
    dict(nsources=3, src_z0=-6.0, src_dz=6.0, nrows=8),
    dict(style=1, nz=16, dv=1.0),
    dict(style=2, nz=16, dv=1.0),
    dict(style=2, x_offset=2, det_offset=1.3, nrows=11, dv=0.7),
])
def test_siddon_matches_brute_force(geom_kw):
    g = small_geom(**geom_kw)
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(60):
        view = int(rng.integers(g.nviews))
        col, row = int(rng.integers(g.ncols)), int(rng.integers(g.nrows))
        src = int(rng.integers(g.nsources))
        x1, x2 = ray_endpoints(g, view, col, row, src)
        ref = brute_force_segments(g, x1, x2)
        got = oracle_segments(g, view, col, row, src)
        assert set(ref) == set(got), (view, col, row)
        for k in ref:
            assert got[k] == pytest.approx(ref[k], abs=1e-9)
        checked += len(ref) > 0
    assert checked > 20


def chord_length(x1, x2, lo, hi):
    d = x2 - x1
    a0, a1 = -np.inf, np.inf
    for a in range(3):
        if abs(d[a]) < 1e-15:
            if not lo[a] <= x1[a] <= hi[a]:
                return 0.0
            continue
        t0, t1 = (lo[a] - x1[a]) / d[a], (hi[a] - x1[a]) / d[a]
        a0, a1 = max(a0, min(t0, t1)), min(a1, max(t0, t1))
    a0, a1 = max(a0, 0.0), min(a1, 1.0)
    return max(0.0, a1 - a0) * np.linalg.norm(d)


@pytest.mark.parametrize("style", [0, 1, 2])
def test_uniform_volume_gives_chord_length(style):
    g = small_geom(style=style)
    vol = np.ones(g.vol_shape)
    proj = orc.forward(g, vol, prec=64)
    lo = np.array([-g.nx / 2, -g.ny / 2, -g.nz / 2]) * g.vx
    hi = -lo
    for view in range(g.nviews):
        for row in range(g.nrows):
            for col in range(g.ncols):
                x1, x2 = ray_endpoints(g, view, col, row)
                ch = chord_length(x1, x2, lo, hi)
                expect = ch if ch >= 1e-6 * g.vx else 0.0
                assert proj[view, row, col] == pytest.approx(expect, rel=1e-10, abs=1e-9)


def test_fp32_forward_close_to_fp64():
    g = adjoint_64().with_views(12)
    vol = phantom(g, seed=3).numpy()
    p32 = orc.forward(g, vol, prec=32)
    p64 = orc.forward(g, vol, prec=64)
    rel = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
    assert rel < 2e-6


def test_voxel_backprojection_constant_correction():
    """c = kappa on every pixel gives dx = beta*kappa wherever L >= 1e-6 vx."""
    g = small_geom()
    corr = np.full((1, g.nrows, g.ncols), 0.75)
    out = orc.backproject_voxel(g, np.zeros(g.vol_shape), corr, view0=2, nviews=1, beta=0.5, mode=1, prec=64)
    ones = orc.backproject_voxel(g, np.zeros(g.vol_shape), np.ones_like(corr), view0=2, nviews=1, beta=1.0,
                                 mode=0, prec=64)
    hit = ones > 1e-6 * g.vx
    assert hit.sum() > 100
    np.testing.assert_allclose(out[hit], 0.5 * 0.75, rtol=1e-12)
    assert np.all(out[~hit] == 0)
    zero = orc.backproject_voxel(g, np.full(g.vol_shape, 3.0), np.zeros_like(corr), view0=2, nviews=1,
                                 beta=1.0, mode=1, prec=64)
    assert np.all(zero == 3.0)


@pytest.mark.parametrize("style", [0, 1, 2])
def test_voxel_chords_equal_siddon_lengths(style):
    """The voxel-driven chord sum with c = 1 (W) equals A^T 1 built from the
    independent Siddon traversal, for every voxel inside the reconstruction
    cylinder (the mask and the 1e-6 cut are the only differences).  Fan and
    parallel beam project each voxel onto its own row (rows at the slice
    centres, as the reference sets Z_prj = ZETA and vz = du)."""
    g = small_geom(nx=10, ny=10, nz=6) if style == 0 else small_geom(nx=10, ny=10, nz=16, style=style)
    for view in (0, 3):
        W = orc.backproject_voxel(g, np.zeros(g.vol_shape), np.ones((1, g.nrows, g.ncols)), view0=view,
                                  nviews=1, beta=1.0, mode=0, prec=64)
        proj = np.zeros((g.nviews, g.nrows, g.ncols))
        proj[view] = 1.0
        at1 = orc.adjoint(g, proj, prec=64)
        ii, jj = np.meshgrid(np.arange(g.nx), np.arange(g.ny), indexing="xy")
        hx, hy = g.nx * 0.5 - 0.5, g.ny * 0.5 - 0.5
        inside = ((ii - hx) ** 2 + (jj - hy) ** 2) < hx * hy
        mask = np.broadcast_to(inside, g.vol_shape)
        np.testing.assert_allclose(W[mask], at1[mask], rtol=0, atol=1e-5)


def test_sart_inverse_crime_fixed_point():
    g = small_geom()
    x = phantom(g, seed=5, scale=0.15).numpy().astype(np.float64)
    b = orc.forward(g, x, prec=64)
    y = orc.sart_sweep(g, x, b, beta=1.0, subset=1, prec=64)
    assert np.array_equal(x, y)


@pytest.mark.parametrize("subset,order", [(1, "ascending"), (3, "interleaved"), (2, "ascending")])
def test_sart_reduces_residual(subset, order):
    g = small_geom()
    x = phantom(g, seed=5, scale=0.15).numpy().astype(np.float64)
    x[2:5, 3:7, 4:9] += 0.02
    b = orc.forward(g, x, prec=64)
    v = np.zeros_like(x)
    res = [np.linalg.norm(orc.forward(g, v, prec=64) - b)]
    for _ in range(3):
        v = orc.sart_sweep(g, v, b, beta=1.0, subset=subset, order=order, prec=64)
        res.append(np.linalg.norm(orc.forward(g, v, prec=64) - b))
    assert res[-1] < 0.5 * res[0]
    assert all(r1 <= r0 * 1.0001 for r0, r1 in zip(res, res[1:]))


def test_sart_subset_one_is_view_by_view():
    """subset = 1 is the reference order: correction of view j, then its update."""
    g = small_geom()
    x = phantom(g, seed=2, scale=0.15).numpy().astype(np.float32)
    b = orc.forward(g, x, prec=32) * 1.05
    v = np.zeros_like(x)
    for j in range(g.nviews):
        c = orc.sart_correction(g, v, b[j:j + 1], view0=j, nviews=1, prec=32)
        v = orc.backproject_voxel(g, v, c, view0=j, nviews=1, beta=0.8, mode=1, prec=32)
    w = orc.sart_sweep(g, np.zeros_like(x), b, beta=0.8, subset=1, prec=32)
    assert np.array_equal(v, w)


@pytest.mark.parametrize("prec,tol", [(64, 1e-12), (32, 1e-6)])
def test_adjoint_dot_product(prec, tol):
    g = adjoint_64().with_views(9)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(g.vol_shape)
    y = rng.standard_normal((g.nviews, g.nrows, g.ncols))
    ax = orc.forward(g, x, prec=prec).astype(np.float64)
    aty = orc.adjoint(g, y, prec=prec).astype(np.float64)
    lhs = float(np.dot(ax.ravel(), y.astype(orc._np_dtype(prec)).astype(np.float64).ravel()))
    rhs = float(np.dot(x.astype(orc._np_dtype(prec)).astype(np.float64).ravel(), aty.ravel()))
    assert abs(lhs - rhs) <= tol * abs(lhs)


def test_fista_momentum_closed_form():
    rng = np.random.default_rng(0)
    x, xp = rng.standard_normal(1000), rng.standard_normal(1000)
    t = 1.0
    for _ in range(4):
        tn = (1 + math.sqrt(1 + 4 * t * t)) / 2
        g = (t - 1) / tn
        y, p = orc.fista_momentum(x, xp, g, prec=64)
        np.testing.assert_allclose(y, x + g * (x - xp), rtol=1e-15)
        np.testing.assert_array_equal(p, x)
        t = tn
    # first step: gamma_1 = 0 so y_1 = x_1
    y, _ = orc.fista_momentum(x, xp, 0.0, prec=64)
    np.testing.assert_array_equal(y, x)


def test_fgp_difference_pair_is_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5, 6, 7))
    p = [rng.standard_normal((5, 6, 7)) for _ in range(3)]
    Dx = orc.tv_D(x)
    # duals on the last plane of their axis are structurally zero
    p[0][:, :, -1] = 0
    p[1][:, -1, :] = 0
    p[2][-1, :, :] = 0
    lhs = sum(float((a * b).sum()) for a, b in zip(Dx, p))
    rhs = float((x * orc.tv_Dt(*p)).sum())
    assert lhs == pytest.approx(rhs, rel=1e-13)


def fgp_objective(x, b, lam):
    r, s, t = orc.tv_D(x)
    return float(((x - b) ** 2).sum() + 2 * lam * np.sqrt(r * r + s * s + t * t).sum())


def test_fgp_properties():
    g = small_geom()
    b = phantom(g, seed=4, scale=0.15).numpy().astype(np.float64)
    b += np.random.default_rng(1).normal(0, 0.003, b.shape)
    # constant input is a fixed point
    c = np.full((6, 7, 8), 0.3)
    np.testing.assert_allclose(orc.tv_fgp(c, 0.01, 10, prec=64), c, atol=1e-15)
    # lambda -> 0 returns b
    np.testing.assert_allclose(orc.tv_fgp(b, 1e-9, 5, prec=64), b, atol=1e-7)
    # the denoised volume lowers ||x-b||^2 + 2 lambda TV(x), more so with more iterations
    lam = 0.004
    o0 = fgp_objective(b, b, lam)
    o5 = fgp_objective(orc.tv_fgp(b, lam, 5, prec=64), b, lam)
    o40 = fgp_objective(orc.tv_fgp(b, lam, 40, prec=64), b, lam)
    assert o40 < o5 < o0
    # FP32 instantiation agrees with FP64
    x32 = orc.tv_fgp(b, lam, 20, prec=32)
    x64 = orc.tv_fgp(b, lam, 20, prec=64)
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-5


def test_tv_norm_closed_forms():
    M, N, Z = 7, 6, 5
    a = 0.37
    ramp = np.broadcast_to(a * np.arange(M, dtype=np.float64), (Z, N, M)).copy()
    assert orc.tv_norm(ramp, prec=64) == pytest.approx(a * (M - 1) * N * Z, rel=1e-13)
    # isotropic: ramp along x and y with equal slope gives sqrt(2) a on the interior
    xy = a * (np.arange(M)[None, None, :] + np.arange(N)[None, :, None]) * np.ones((Z, 1, 1))
    expect = a * (math.sqrt(2) * (M - 1) * (N - 1) + (M - 1) + (N - 1)) * Z
    assert orc.tv_norm(xy, prec=64) == pytest.approx(expect, rel=1e-13)
    u = np.random.default_rng(0).standard_normal(500)
    v = np.random.default_rng(1).standard_normal(500)
    assert orc.residual_norm2(u, v, prec=64) == pytest.approx(float(((u - v) ** 2).sum()), rel=1e-13)


def test_tv_eps_gradient_matches_central_differences():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 5, 6)) * 0.1
    eps, h = 1e-8, 1e-6
    g = orc.tv_eps_grad(x, eps)
    for _ in range(25):
        k, j, i = rng.integers(4), rng.integers(5), rng.integers(6)
        xp, xm = x.copy(), x.copy()
        xp[k, j, i] += h
        xm[k, j, i] -= h
        fd = (orc.tv_eps(xp, eps) - orc.tv_eps(xm, eps)) / (2 * h)
        assert g[k, j, i] == pytest.approx(fd, rel=1e-6, abs=1e-7)
    # 1-D ramp closed form: TV_eps = (n-1) sqrt(eps + a^2) + sqrt(eps) per line
    a, n = 0.25, 6
    line = np.broadcast_to(a * np.arange(n, dtype=np.float64), (1, 1, n)).copy()
    assert orc.tv_eps(line, eps) == pytest.approx((n - 1) * math.sqrt(eps + a * a) + math.sqrt(eps), rel=1e-13)


@pytest.mark.parametrize("n", [8, 64, 512, 12, 96])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(orc.fft(x, -1), np.fft.fft(x), rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(orc.fft(x, +1), np.fft.ifft(x) * n, rtol=1e-12, atol=1e-10)


def test_ramp_taps_and_filter_match_numpy():
    g = Geometry(nx=8, ny=8, nz=8, vx=1.0, ncols=32, nrows=4, du=0.8, nviews=3, dso=100.0, dod=-60.0)
    pad = 2 * g.ncols
    fmax = 1.0 / 2.0 / (g.du * g.dso / (g.dso - g.dod))
    k = np.fft.fftfreq(pad) * pad
    np.testing.assert_allclose(orc.ramp_taps(g, "ramp"), 0.5 * np.abs(k) * fmax / (pad / 2), rtol=1e-6)
    hann = 0.5 * (1 + np.cos(2 * np.pi * k / pad))
    np.testing.assert_allclose(orc.ramp_taps(g, "hann"), 0.5 * np.abs(k) * fmax / (pad / 2) * hann, rtol=1e-6)
    rng = np.random.default_rng(9)
    p = rng.standard_normal((g.nviews, g.nrows, g.ncols))
    for w in ("ramp", "hann"):
        H = orc.ramp_taps(g, w)
        padded = np.zeros((g.nviews, g.nrows, pad))
        padded[..., : g.ncols] = p
        ref = np.real(np.fft.ifft(np.fft.fft(padded, axis=-1) * H, axis=-1))[..., : g.ncols]
        np.testing.assert_allclose(orc.ramp_filter(g, p, w, prec=64), ref, rtol=1e-10, atol=1e-12)
    # impulse response of the pure ramp: h[0] > 0, h[n] < 0 for odd n, ~0 for even n != 0
    imp = np.zeros((1, 1, g.ncols))
    imp[0, 0, 10] = 1.0
    g1 = Geometry(**{**g.__dict__, "nrows": 1, "nviews": 1})
    h = orc.ramp_filter(g1, imp, "ramp", prec=64)[0, 0]
    assert h[10] > 0 and h[11] < 0 and h[9] < 0 and abs(h[12]) < 0.05 * h[10]


def test_fdk_weight_formula():
    g = Geometry(nx=8, ny=8, nz=8, vx=1.0, ncols=12, nrows=6, du=0.8, nviews=2, dso=100.0, dod=-60.0)
    p = np.random.default_rng(0).standard_normal((2, 6, 12))
    w = orc.fdk_weight(g, p, prec=64)
    u = (-0.8 * (12 / 2 - 0.5) + 0.8 * np.arange(12)) * 100 / 160
    v = (-0.8 * (6 / 2 - 0.5) + 0.8 * np.arange(6)) * 100 / 160
    W = 100 / np.sqrt(100 ** 2 + u[None, :] ** 2 + v[:, None] ** 2)
    np.testing.assert_allclose(w, p * W[None], rtol=1e-13)


@pytest.mark.slow
def test_fdk_reconstructs_transposed_object():
    """The reference FDK (+theta, depth s = -x sin + y cos) reconstructs the
    forward model's object with x and y exchanged."""
    g = Geometry(nx=32, ny=32, nz=16, vx=1.0, ncols=64, nrows=32, du=1.0, nviews=120, dso=300.0, dod=-150.0)
    x = np.zeros(g.vol_shape)
    x[4:12, 6:10, 18:28] = 0.02      # asymmetric block: y in 6..9, x in 18..27
    x[6:10, 20:26, 5:9] = 0.01
    proj = orc.forward(g, x, prec=64)
    rec = orc.fdk(g, proj, "ramp", prec=64)
    c_same = np.corrcoef(rec.ravel(), x.ravel())[0, 1]
    c_swap = np.corrcoef(rec.ravel(), x.transpose(0, 2, 1).ravel())[0, 1]
    assert c_swap > 0.8 and c_swap > c_same + 0.3


MS_FDK = dict(nx=12, ny=12, nz=10, vx=1.0, ncols=20, nrows=14, du=1.3, nviews=6, dso=60.0, dod=-30.0,
              nsources=3, src_z0=-4.0, src_dz=4.0)


@pytest.mark.parametrize("mode", ["inverse_square", "nearest", "equal"])
def test_multisource_fdk_one_source_is_fdk(mode):
    """With one source at height 0 the three source combinations reduce to the
    single-source FDK backprojection (the weights normalise to 1; the nearest
    source is the only one)."""
    g = Geometry(**{**MS_FDK, "nsources": 1, "src_z0": 0.0, "src_dz": 0.0})
    f = np.random.default_rng(3).standard_normal((g.nviews, g.nrows, g.ncols))
    v0 = np.zeros(g.vol_shape)
    ref = orc.fdk_backproject(g, f, v0, prec=64)
    got = orc.fdk_backproject_ms(g, f, v0, mode=mode, prec=64)
    if mode == "inverse_square":
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-15)
    else:
        assert np.array_equal(got, ref)


def test_multisource_fdk_combinations():
    """Source combinations pinned by invariants on one view: "equal" is linear
    in the per-source images; with all images 1 "inverse_square" gives the
    largest single-source value (the weights normalise) wherever a source
    reaches the voxel; with images c_i it is a convex combination of the
    reaching sources' c_i; "nearest" returns c of the reaching source
    nearest the voxel in z."""
    g = Geometry(**{**MS_FDK, "nviews": 1})
    shape = (1, g.nsources, g.nrows, g.ncols)
    zero = np.zeros(g.vol_shape)
    c = np.array([1.0, 10.0, 100.0])
    single = []
    for i in range(g.nsources):
        f = np.zeros(shape)
        f[0, i] = 1.0
        single.append(orc.fdk_backproject_ms(g, f, zero, mode="equal", prec=64))   # t1*k where source i reaches
    single = np.stack(single)
    reach = single > 0
    assert reach.any(axis=0).mean() > 0.5 and (reach.sum(axis=0) >= 2).any()
    fc = np.broadcast_to(c[None, :, None, None], shape).copy()
    equal = orc.fdk_backproject_ms(g, fc, zero, mode="equal", prec=64)
    np.testing.assert_allclose(equal, np.tensordot(c, single, axes=1), rtol=1e-13)
    ones = orc.fdk_backproject_ms(g, np.ones(shape), zero, mode="inverse_square", prec=64)
    np.testing.assert_allclose(ones, single.max(axis=0), rtol=1e-13)
    inv = orc.fdk_backproject_ms(g, fc, zero, mode="inverse_square", prec=64)
    tk = single.max(axis=0)
    lo = np.where(reach, c[:, None, None, None], np.inf).min(axis=0) * tk
    hi = np.where(reach, c[:, None, None, None], -np.inf).max(axis=0) * tk
    m = reach.any(axis=0)
    assert np.all(inv[m] >= lo[m] * (1 - 1e-12)) and np.all(inv[m] <= hi[m] * (1 + 1e-12))
    assert np.all(inv[~m] == 0)
    near = orc.fdk_backproject_ms(g, fc, zero, mode="nearest", prec=64)
    zc = -g.nz / 2 + 0.5 + np.arange(g.nz)                       # voxel-centre heights (vz = 1)
    zs = g.src_z0 + g.src_dz * np.arange(g.nsources)
    dist = np.abs(zs[:, None] - zc[None, :])[:, :, None, None]  # [source, z, 1, 1]
    dist = np.where(reach, dist, np.inf)
    pick = np.argmin(dist, axis=0)
    want = np.where(m, c[pick] * tk, 0.0)
    np.testing.assert_allclose(near, want, rtol=1e-13)
