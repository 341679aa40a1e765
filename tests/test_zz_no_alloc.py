"""ct_forward_noalloc, the allocation-free variant of ct_forward, through the C
ABI: two calls in a row (the second on a smaller geometry and volume) each
match ct_forward on the same inputs, to the projection bar (1e-5 relative L2
in FP32; exactly for fewer than four images, which take the same kernel;
1e-12 in FP64), and a geometry past the scratch limit is refused with a clear
error.  Runs alone or after the other tests."""
import ctypes

import numpy as np
import pytest
import torch

from gpu_amd import core
from gpu_amd.geometry import Geometry, plain_sart_128
from gpu_amd.phantom import phantom

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.double().cpu().numpy()
    b = b.double().cpu().numpy()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def noalloc(ctx, g, vol, proj, view0, nviews):
    lib = core.lib()
    fn = lib.ct_forward_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.POINTER(core.CtGeometry), ctypes.c_void_p, ctypes.c_int32,
                   ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    fn.restype = ctypes.c_int32
    gc = g.to_c()
    ctx._stream()
    return fn(ctx._h, ctypes.byref(gc), ctypes.c_void_p(vol.data_ptr()), core.dtype_code(vol),
              ctypes.c_void_p(proj.data_ptr()), core.dtype_code(proj), view0, nviews)


def test_forward_noalloc_matches_forward(ctx):
    big = plain_sart_128().with_views(12)
    small = Geometry(nx=40, ny=36, nz=24, vx=1.3, ncols=60, nrows=44, du=1.2, nviews=7, dso=200.0, dod=-120.0,
                     angle_step_deg=23.0, x_offset=2, det_offset=1.5)
    cases = [(big, torch.float32, 2, 9, 1e-5), (small, torch.float32, 1, 3, 0.0),
             (big, torch.float64, 0, 5, 1e-12), (small, torch.float64, 4, 3, 1e-12)]
    for g, dt, v0, nv, tol in cases:
        x = phantom(g, seed=3, scale=0.6 if g is big else 0.3, device="cuda").to(dt)
        got = torch.full((nv, g.nrows, g.ncols), -1.0, device="cuda", dtype=dt)
        st = noalloc(ctx, g, x, got, v0, nv)
        assert st == 0, ctx._h and core.lib().ct_last_error(ctx._h)
        ref = ctx.forward(g, x, view0=v0, nviews=nv)
        torch.cuda.synchronize()
        r = rel(got, ref)
        print(f"{g.nx}^3-ish {dt} views {v0}+{nv}: rel L2 {r:.3e}")
        if tol == 0.0:
            assert torch.equal(got, ref)
        else:
            assert r <= tol
        assert float(got.abs().max()) > 0


def test_forward_noalloc_rejects_too_many_views(ctx):
    g = Geometry(nx=8, ny=8, nz=8, vx=1.0, ncols=8, nrows=8, du=1.0, nviews=16385, dso=100.0, dod=-50.0)
    x = torch.zeros(g.vol_shape, device="cuda")
    p = torch.zeros((1, g.nrows, g.ncols), device="cuda")
    st = noalloc(ctx, g, x, p, 0, 1)
    assert st != 0
    msg = core.lib().ct_last_error(ctx._h).decode()
    assert "16384" in msg, msg
