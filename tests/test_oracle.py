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


# ---------------------------------------------------------------- FDK pins

def cylinder_projections(g, radius, height, mu):
    """Analytic line integrals (FP64) of a centred uniform cylinder
    x^2 + y^2 <= radius^2, |z| <= height/2 of attenuation mu, on the forward
    model's rays (detector pixel -> source)."""
    out = np.zeros((g.nviews, g.nrows, g.ncols))
    u = -g.du * (g.ncols / 2 - 0.5) + np.arange(g.ncols) * g.du
    dv = g.dv or g.du
    v = -dv * (g.nrows / 2 - 0.5) + np.arange(g.nrows) * dv
    for j in range(g.nviews):
        s, c = orc.view_angle(g, j, prec=64)
        x1x, x1y = g.dod * c - u * s, g.dod * s + u * c
        dx, dy = g.dso * c - x1x, g.dso * s - x1y
        a = dx * dx + dy * dy
        b = 2 * (x1x * dx + x1y * dy)
        disc = b * b - 4 * a * (x1x ** 2 + x1y ** 2 - radius ** 2)
        hit = disc > 0
        sq = np.sqrt(np.where(hit, disc, 0.0))
        a0, a1 = (-b - sq) / (2 * a), (-b + sq) / (2 * a)
        dz = -v[:, None]
        t0, t1 = (-height / 2 - v[:, None]) / dz, (height / 2 - v[:, None]) / dz
        lo = np.maximum(a0[None, :], np.minimum(t0, t1))
        hi = np.minimum(a1[None, :], np.maximum(t0, t1))
        out[j] = np.where(hit[None, :], np.maximum(hi - lo, 0.0) * np.sqrt(a[None, :] + dz * dz), 0.0) * mu
    return out


@pytest.mark.parametrize("dso,dod", [(100.0, -60.0), (1000.0, -500.0)])
def test_fdk_uniform_cylinder_amplitude(dso, dod):
    """FDK (V63/host_FDK.cpp:4-57) of analytic projections of a uniform
    cylinder over 360 views returns mu, flat, in the central slices: the mean
    inside 12 mm is within 2% of mu (the discrete ramp's known DC deficit is
    ~1.4%) and every 3 mm ring is within 0.4% of that mean.  A dropped or
    doubled 0.5 in the ramp (V63/FDK/FFT_1d_GPU.h:266-267) or a wrong view
    weight (V63/FDK/kernel_FDK.cu:60) moves the mean by 2x or more; a wrong
    power of DSO/(DSO - s) (:36) tilts the rings by 1.6-3% at DSO = 100 mm,
    against 0.23% with the right one."""
    g = Geometry(nx=48, ny=48, nz=4, vx=1.0, ncols=160, nrows=48, du=0.6 * (dso - dod) / dso, nviews=360,
                 dso=dso, dod=dod)
    mu = 0.02
    rec = orc.fdk(g, cylinder_projections(g, 18.0, 60.0, mu), "ramp", prec=64)
    mid = rec[1:3]
    yy, xx = np.meshgrid(np.arange(48) - 23.5, np.arange(48) - 23.5, indexing="ij")
    r = np.hypot(xx, yy)
    mean = mid[:, r < 12].mean()
    rings = np.array([mid[:, (r >= a) & (r < a + 3)].mean() for a in (0, 3, 6, 9, 12)])
    print(f"DSO {dso}: mean/mu {mean / mu:.4f}, rings/mean {np.round(rings / mean, 4)}")
    assert abs(mean / mu - 1) < 0.02
    assert np.all(np.abs(rings / mean - 1) < 0.004)


def test_fdk_backprojection_affine_view():
    """One filtered view that is affine in (row, column), q = a + b col + c row,
    backprojects to DSO^2/(DSO - s)^2 (a + b p + c ksi) 2 pi / N_views exactly
    (V63/FDK/kernel_FDK.cu:33-60): bilinear interpolation reproduces an affine
    function, so the test fixes the stencil's orientation (b != c catches
    swapped row and column fractions), the distance weight and the view factor
    dtheta * 360 / (us_rate * N_views) = 2 pi / N_views.  The detector
    coordinates p and ksi are computed here from the geometry's definition:
    voxel centre, +theta rotation, magnification to the detector plane."""
    g = Geometry(nx=14, ny=12, nz=10, vx=1.1, ncols=40, nrows=30, du=0.9, nviews=7, dso=80.0, dod=-50.0,
                 angle_step_deg=23.0, angle0_deg=11.0)
    a, b, c = 0.7, 0.013, -0.029
    rows, cols = np.meshgrid(np.arange(g.nrows), np.arange(g.ncols), indexing="ij")
    for view in (0, 3, 6):
        filt = np.zeros((g.nviews, g.nrows, g.ncols))
        filt[view] = a + b * cols + c * rows
        got = orc.fdk_backproject(g, filt, np.zeros(g.vol_shape), prec=64)
        th = math.radians(view * 23.0 + 11.0)
        sn, cs = math.sin(th), math.cos(th)
        zc = (np.arange(g.nz) - g.nz / 2 + 0.5) * g.vx
        yc = (np.arange(g.ny) - g.ny / 2 + 0.5) * g.vx
        xc = (np.arange(g.nx) - g.nx / 2 + 0.5) * g.vx
        Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
        t = X * cs + Y * sn
        s = -X * sn + Y * cs
        mag = (g.dso - g.dod) / (g.dso - s)
        p = (t * mag + g.du * (g.ncols / 2 - 0.5)) / g.du
        ksi = (Z * mag + g.du * (g.nrows / 2 - 0.5)) / g.du
        inside = (ksi >= 0) & (ksi + 1 < g.nrows) & (p >= 0) & (p + 1 < g.ncols)
        want = np.where(inside, (g.dso / (g.dso - s)) ** 2 * (a + b * p + c * ksi) * 2 * math.pi / g.nviews, 0.0)
        assert inside.mean() > 0.5
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------- FGP pins

def np_D(x):
    """Forward differences x_v - x_{v+e}, zero on the last plane (numpy, FP64)."""
    r, s, t = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    r[:, :, :-1] = x[:, :, :-1] - x[:, :, 1:]
    s[:, :-1, :] = x[:, :-1, :] - x[:, 1:, :]
    t[:-1, :, :] = x[:-1, :, :] - x[1:, :, :]
    return r, s, t


def np_Dt(r, s, t):
    out = r.copy() + s + t
    out[:, :, 1:] -= r[:, :, :-1]
    out[:, 1:, :] -= s[:, :-1, :]
    out[1:, :, :] -= t[:-1, :, :]
    return out


def unit_ball(r, s, t):
    n = np.maximum(1.0, np.sqrt(r * r + s * s + t * t))
    return r / n, s / n, t / n


@pytest.mark.parametrize("lam", [0.5, 1e-3])
def test_fgp_first_iteration_closed_form(lam):
    """From zero duals one FGP step (V63/kernel_FGP_Denoise_GPUx7.cu:2-88,
    V63/host_FGP_Denoise_GPUx7.c:64-86) is the projected gradient step of the
    dual of min ||x - b||^2 + 2 lam TV(x) with step 1/(12 lam):
    x1 = b - lam D^T P(D b / (12 lam)), P the per-voxel projection onto the
    unit ball.  lam = 0.5 keeps every |Db|/(12 lam) < 1 (x1 = b - D^T D b / 12,
    a Laplacian step independent of lam), lam = 1e-3 saturates the
    projection.  A wrong step constant, lambda scaling or projection fails."""
    rng = np.random.default_rng(5)
    b = rng.standard_normal((6, 7, 8)) * 0.1
    x1 = orc.tv_fgp(b, lam, 1, prec=64)
    r, s, t = np_D(b)
    want = b - lam * np_Dt(*unit_ball(r / (12 * lam), s / (12 * lam), t / (12 * lam)))
    np.testing.assert_allclose(x1, want, rtol=1e-12, atol=1e-14)
    if lam == 0.5:
        np.testing.assert_allclose(x1, b - np_Dt(r, s, t) / 12, rtol=1e-12, atol=1e-14)


def prox_tv_pdhg(b, lam, iters=8000):
    """argmin_x ||x - b||^2 + 2 lam sum_v |(Dx)_v| by the primal-dual method of
    Chambolle and Pock (2011, algorithm 2: G strongly convex, accelerated),
    an independent solver for the FGP fixed point."""
    x = b.copy()
    xb = x.copy()
    y = [np.zeros_like(b) for _ in range(3)]
    L2 = 12.0                    # ||D||^2 <= 12 in 3-D
    tau, sigma, gam = 1.0, 1.0 / 12.0, 2.0
    for _ in range(iters):
        d = np_D(xb)
        y = [yi + sigma * di for yi, di in zip(y, d)]
        n = np.maximum(1.0, np.sqrt(y[0] ** 2 + y[1] ** 2 + y[2] ** 2) / (2 * lam))
        y = [yi / n for yi in y]
        xn = (x - tau * np_Dt(*y) + 2 * tau * b) / (1 + 2 * tau)
        theta = 1 / math.sqrt(1 + 2 * gam * tau)
        tau, sigma = theta * tau, sigma / theta
        xb = xn + theta * (xn - x)
        x = xn
    assert tau * sigma * L2 <= 1.0 + 1e-9
    return x


def test_fgp_converges_to_tv_prox():
    """Run long enough, FGP returns argmin ||x - b||^2 + 2 lam TV(x) with the
    isotropic forward-difference TV of the reference (zero on the last
    planes): it agrees with an independent primal-dual solver to 1e-4 of the
    volume (both converge like 1/k; the solvers' own distance to the
    minimiser is ~2e-5) and reaches the same objective to 1e-7.  A different
    lambda convention (||x - b||^2 + lam TV, or a lost factor in the step)
    moves the fixed point by more than 1% here.  The FP32 instantiation
    follows FP64."""
    rng = np.random.default_rng(11)
    b = np.zeros((8, 9, 10))
    b[2:6, 2:7, 3:8] = 1.0
    b += rng.normal(0, 0.1, b.shape)
    lam = 0.05
    ref = prox_tv_pdhg(b, lam)
    got = orc.tv_fgp(b, lam, 3000, prec=64)
    obj = lambda x: float(((x - b) ** 2).sum() + 2 * lam * np.sqrt(sum(d * d for d in np_D(x))).sum())
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"FGP(3000) vs primal-dual: rel {err:.2e}, objectives {obj(got):.12f} {obj(ref):.12f}")
    assert err < 1e-4
    assert abs(obj(got) - obj(ref)) <= 1e-7 * obj(ref)
    assert np.linalg.norm(orc.tv_fgp(b, 2 * lam, 3000, prec=64) - ref) / np.linalg.norm(ref) > 1e-2
    assert np.linalg.norm(orc.tv_fgp(b, lam / 2, 3000, prec=64) - ref) / np.linalg.norm(ref) > 1e-2
    x32 = orc.tv_fgp(b, lam, 200, prec=32)
    x64 = orc.tv_fgp(b, lam, 200, prec=64)
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-5


def test_fgp_duals_in_unit_ball_and_acceleration():
    """After every iteration the projected duals (o, p, q) lie in the unit
    ball per voxel (V63/kernel_FGP_Denoise_GPUx7.cu:67-85), and the momentum
    makes FGP converge faster than the plain projected gradient (the same
    iteration with gamma = 0, emulated from single steps): after 30
    iterations FGP's objective gap is below a third of the gradient's."""
    rng = np.random.default_rng(2)
    b = rng.standard_normal((5, 6, 7)) * 0.3
    lam = 0.08
    for n in (1, 2, 3, 5, 8, 13, 30):
        _, duals = orc.tv_fgp_duals(b, lam, n, prec=64)
        nrm = np.sqrt((duals ** 2).sum(axis=0))
        assert nrm.max() <= 1.0 + 1e-12, n
        assert nrm.max() > 0.99        # the projection is active for this lambda
    ref = prox_tv_pdhg(b, lam)
    obj = lambda x: float(((x - b) ** 2).sum() + 2 * lam * np.sqrt(sum(d * d for d in np_D(x))).sum())
    gap_fgp = obj(orc.tv_fgp(b, lam, 30, prec=64)) - obj(ref)
    # plain projected gradient on the dual: p <- P(p + D(b - lam D^T p)/(12 lam))
    p = [np.zeros_like(b) for _ in range(3)]
    for _ in range(30):
        L = b - lam * np_Dt(*p)
        d = np_D(L)
        p = list(unit_ball(*[pi + di / (12 * lam) for pi, di in zip(p, d)]))
    gap_gp = obj(b - lam * np_Dt(*p)) - obj(ref)
    print(f"objective gap after 30: FGP {gap_fgp:.3e}, projected gradient {gap_gp:.3e}")
    assert 0 <= gap_fgp < gap_gp / 3


def cpu_fdnoise(b, lam, niter):
    """numpy restatement of the reference's CPU FGP, F_Dnoise
    (V63/host_FGP_Denoise_CPU.h:242-347) with F_OperatorL (:57-120),
    F_OperatorLT (:134-153) and F_ProjectionP (:157-238) on its compact dual
    arrays X: Z*Y*(X-1), Y: Z*(Y-1)*X, Z: (Z-1)*Y*X, including its
    out-of-step reads on the last x, y and z planes (reads past the end of an
    array return 0 here)."""
    NZ, NY, NX = b.shape
    nX, nY, nZ = NZ * NY * (NX - 1), NZ * (NY - 1) * NX, (NZ - 1) * NY * NX
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")

    def rd(a, idx):
        a1 = np.concatenate([a, np.zeros(4 * NX * NY + 8)])
        return np.where(idx >= 0, a1[np.clip(idx, 0, a1.size - 1)], 0.0)

    def opL(X, Y, Z):
        xi = (k * NY + j) * (NX - 1) + i
        yi = (k * (NY - 1) + j) * NX + i
        zi = (k * NY + j) * NX + i
        A = rd(X, xi) - np.where(i > 0, rd(X, xi - 1), 0)
        A += rd(Y, yi) - np.where(j > 0, rd(Y, ((k * (NY - 1) + (j - 1)) * NX + i)), 0)
        A += rd(Z, zi) - np.where(k > 0, rd(Z, (((k - 1) * NY + j) * NX + i)), 0)
        return A

    def opLT(A):
        X = (A[:, :, :-1] - A[:, :, 1:]).ravel()
        Y = (A[:, :-1, :] - A[:, 1:, :]).ravel()
        Z = (A[:-1, :, :] - A[1:, :, :]).ravel()
        return X, Y, Z

    def projP(X1, Y1, Z1):
        X, Y, Z = np.zeros(nX), np.zeros(nY), np.zeros(nZ)
        for kk in range(NZ):
            for jj in range(NY):
                for ii in range(NX):
                    xi = (kk * NY + jj) * (NX - 1) + ii
                    yi = (kk * (NY - 1) + jj) * NX + ii
                    zi = (kk * NY + jj) * NX + ii
                    hx, hy, hz = ii < NX - 1, jj < NY - 1, kk < NZ - 1
                    q = (X1[xi] ** 2 if hx else 0) + (Y1[yi] ** 2 if hy else 0) + (Z1[zi] ** 2 if hz else 0)
                    den = max(1.0, math.sqrt(q))
                    if hx:
                        X[xi] = X1[xi] / den
                    if hy:
                        Y[yi] = Y1[yi] / den
                    if hz:
                        Z[zi] = Z1[zi] / den
        return X, Y, Z

    X0, Y0, Z0 = np.zeros(nX), np.zeros(nY), np.zeros(nZ)
    Xm, Ym, Zm = np.zeros(nX), np.zeros(nY), np.zeros(nZ)
    X, Y, Z = Xm, Ym, Zm
    t = 1.0
    for _ in range(niter):
        tmp = b - lam * opL(X0, Y0, Z0)
        X1, Y1, Z1 = opLT(tmp)
        X, Y, Z = projP(X0 + X1 / (12 * lam), Y0 + Y1 / (12 * lam), Z0 + Z1 / (12 * lam))
        tp1 = (1 + math.sqrt(1 + 4 * t * t)) / 2
        X0, Y0, Z0 = X + (X - Xm) * (t - 1) / tp1, Y + (Y - Ym) * (t - 1) / tp1, Z + (Z - Zm) * (t - 1) / tp1
        Xm, Ym, Zm = X, Y, Z
        t = tp1
    return b - lam * opL(X, Y, Z)


def test_fgp_gpux7_equals_cpu_fgp_in_the_interior():
    """The oracle follows the GPUx7 kernels.  The reference's own CPU FGP
    (F_Dnoise) is the same algorithm on compact dual arrays, except that its
    divergence reads out of step on the last x, y and z planes; those reads
    reach one more voxel inward per iteration.  Away from the last planes the
    two agree to rounding, which ties the oracle to the reference's second
    statement of the algorithm; on the last planes they differ."""
    rng = np.random.default_rng(8)
    b = rng.standard_normal((12, 13, 14)) * 0.2
    lam, n = 0.03, 3
    cpu = cpu_fdnoise(b, lam, n)
    got = orc.tv_fgp(b, lam, n, prec=64)
    m = 2 * n + 2
    inner = (slice(0, 12 - m), slice(0, 13 - m), slice(0, 14 - m))
    np.testing.assert_allclose(got[inner], cpu[inner], rtol=1e-12, atol=1e-14)
    assert np.abs(got - cpu).max() > 1e-6


# ------------------------------------------------ multi-source voxel backprojection pins

def test_multisource_voxel_chords_and_threshold():
    """Multi-source voxel-driven backprojection (TB/voxel/kernel_IterativeRecon_TBCT_voxel.cu:546-775):
    with c = 1 on every image of a view, W summed over the view's sources
    equals A^T 1 of that view built from the independent Siddon traversal at
    every voxel inside the reconstruction cylinder, so the source culling
    (:618-627) drops no source whose rays reach the voxel.  With c = kappa the
    update is beta*kappa where L > 1e-3 vx and nothing elsewhere (:772)."""
    g = small_geom(nx=10, ny=10, nz=12, nsources=4, src_z0=-7.5, src_dz=5.0, nrows=10)
    ii, jj = np.meshgrid(np.arange(g.nx), np.arange(g.ny), indexing="xy")
    hx, hy = g.nx * 0.5 - 0.5, g.ny * 0.5 - 0.5
    mask = np.broadcast_to(((ii - hx) ** 2 + (jj - hy) ** 2) < hx * hy, g.vol_shape)
    for view in (0, 2, 5):
        ones = np.ones((1, g.nsources, g.nrows, g.ncols))
        W = orc.backproject_voxel(g, np.zeros(g.vol_shape), ones, view0=view, nviews=1, beta=1.0, mode=0, prec=64)
        proj = np.zeros((g.nviews, g.nsources, g.nrows, g.ncols))
        proj[view] = 1.0
        at1 = orc.adjoint(g, proj, prec=64)
        thr = at1 > 1e-3 * g.vx
        assert (thr & mask).sum() > 200
        np.testing.assert_allclose(W[mask & thr], at1[mask & thr], rtol=1e-9, atol=1e-9)
        upd = orc.backproject_voxel(g, np.zeros(g.vol_shape), 0.6 * ones, view0=view, nviews=1, beta=0.5,
                                    mode=1, prec=64)
        keep = mask & ((at1 > 1.001e-3 * g.vx) | (at1 < 0.999e-3 * g.vx))
        np.testing.assert_allclose(upd[keep], np.where(thr, 0.3, 0.0)[keep], rtol=1e-12, atol=1e-15)
