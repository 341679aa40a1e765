"""The reference-named wrappers of include/ct_compat.h, called through the C ABI
on host arrays exactly as code written against the reference would call them,
against the CPU oracle.  Geometry is set once per test with
ct_compat_set_geometry (the reference's compile-time constants)."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

from gpu_amd import core
from gpu_amd.geometry import Geometry
from gpu_amd.phantom import phantom
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FP = ctypes.POINTER(ctypes.c_float)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fp(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(FP)


@pytest.fixture()
def compat(ctx):
    lib = core.lib()
    sig = {
        "ct_compat_set_geometry": ([ctypes.POINTER(core.CtGeometry)], ctypes.c_int),
        "ct_compat_last_status": ([], ctypes.c_int),
        "ct_compat_release": ([], None),
        "ct_compat_last_inner_products": ([ctypes.POINTER(ctypes.c_double)] * 2, ctypes.c_int),
        "Forward_3D_ray_driven_siddon": ([FP, FP], None),
        "Forward_3D_ray_driven_siddon_TBCT": ([FP, FP], None),
        "Reconstruction_3D_ray_driven_CBCT": ([FP, FP, ctypes.c_float], None),
        "Reconstruction_3D_ray_driven_TBCT": ([FP, FP, ctypes.c_float], None),
        "CheckMatchedJointOperator": ([FP], None),
        "ct_compat_check_matched_adjoint": ([FP], None),
        "FGP_denoise_GPUx7_exact": ([FP, ctypes.c_float, ctypes.c_int], None),
        "F_Dnoise": ([FP, FP, ctypes.c_float, ctypes.c_int], None),
        "Reconstruction_FDK": ([FP, FP], None),
        "RampFiltering_GPU": ([FP, FP], None),
        "L2_norm_gpu": ([FP, FP], ctypes.c_double),
        "TV_norm_gpu": ([FP], ctypes.c_double),
        "L2_norm_cpu": ([FP, FP, ctypes.c_int], ctypes.c_double),
        "TV_norm_cpu": ([FP], ctypes.c_double),
        "SaveDeviceDataToFile": ([FP, ctypes.c_int, ctypes.c_char_p], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res

    def use(g):
        gc = g.to_c()
        assert lib.ct_compat_set_geometry(ctypes.byref(gc)) == 0
        return lib

    yield use
    lib.ct_compat_release()


G = Geometry(nx=24, ny=20, nz=16, vx=1.0, ncols=40, nrows=28, du=1.1, nviews=9, dso=90.0, dod=-45.0,
             angle_step_deg=37.3)
GMS = Geometry(nx=24, ny=20, nz=16, vx=1.0, ncols=40, nrows=12, du=1.1, nviews=6, dso=90.0, dod=-45.0,
               angle_step_deg=37.3, nsources=4, src_z0=-9.0, src_dz=6.0)


def vol_of(g, seed=18, scale=0.3):
    return np.ascontiguousarray(phantom(g, seed=seed, scale=scale).numpy().astype(np.float32))


def test_check_matched_joint_operator(compat):
    """V63's check (host_IterativeRecon_CBCT.c:131-204): view 0 forward, then
    the voxel-driven backprojector with command 0 into a zeroed volume;
    <A0 x, A0 x> and <x, B0 A0 x> against the oracle's same two steps.  The
    matched variant's two products agree to FP32 rounding, the voxel-driven
    one's differ by the mask and thresholds."""
    lib = compat(G)
    x = vol_of(G)
    lib.CheckMatchedJointOperator(fp(x))
    assert lib.ct_compat_last_status() == 0
    ip1, ip2 = ctypes.c_double(), ctypes.c_double()
    lib.ct_compat_last_inner_products(ctypes.byref(ip1), ctypes.byref(ip2))
    p0 = orc.forward(G, x, prec=32, view0=0, nviews=1)
    bp = orc.backproject_voxel(G, np.zeros_like(x), p0, view0=0, nviews=1, beta=1.0, mode=0, prec=32)
    want1 = float((p0.astype(np.float64) ** 2).sum())
    want2 = float((x.astype(np.float64) * bp).sum())
    print(f"voxel-driven check: {ip1.value:.6f} {ip2.value:.6f} (oracle {want1:.6f} {want2:.6f})")
    assert ip1.value == pytest.approx(want1, rel=1e-5)
    assert ip2.value == pytest.approx(want2, rel=1e-5)
    lib.ct_compat_check_matched_adjoint(fp(x))
    assert lib.ct_compat_last_status() == 0
    lib.ct_compat_last_inner_products(ctypes.byref(ip1), ctypes.byref(ip2))
    assert ip1.value == pytest.approx(want1, rel=1e-5)
    assert ip2.value == pytest.approx(ip1.value, rel=1e-5)
    assert abs(want2 - want1) > 1e-4 * want1     # the voxel-driven B0 is not the transpose


def test_fgp_wrappers(compat):
    """FGP_denoise_GPUx7_exact (V63/host_FGP_Denoise_GPUx7.c:1) in place and
    F_Dnoise (V63/host_FGP_Denoise_CPU.h:242, GPUx7 semantics) into a second
    array, against the oracle's FP32 FGP."""
    lib = compat(G)
    b = vol_of(G, seed=4)
    b += np.random.default_rng(1).normal(0, 0.003, b.shape).astype(np.float32)
    ref = orc.tv_fgp(b, 0.004, 15, prec=32)
    a = b.copy()
    lib.FGP_denoise_GPUx7_exact(fp(a), 0.004, 15)
    assert lib.ct_compat_last_status() == 0
    assert rel(a, ref) <= 1e-5
    clean = np.zeros_like(b)
    noisy = b.copy()
    lib.F_Dnoise(fp(noisy), fp(clean), 0.004, 15)
    assert lib.ct_compat_last_status() == 0
    assert np.array_equal(noisy, b)
    assert rel(clean, ref) <= 1e-5


def test_fdk_and_ramp_wrappers(compat):
    """Reconstruction_FDK (V63/host_FDK.cpp:4-57) accumulates the pure-ramp
    FDK into the caller's volume; RampFiltering_GPU (V63/FDK/FFT_1d_GPU.h:241)
    filters every view out of place and in place."""
    g = Geometry(nx=32, ny=32, nz=16, vx=1.0, ncols=64, nrows=24, du=1.0, nviews=20, dso=300.0, dod=-150.0)
    lib = compat(g)
    x = vol_of(g, seed=9, scale=0.4)
    p = np.ascontiguousarray(orc.forward(g, x, prec=32))
    v0 = vol_of(g, seed=2, scale=0.2)
    vol = v0.copy()
    lib.Reconstruction_FDK(fp(vol), fp(p))
    assert lib.ct_compat_last_status() == 0
    ref = orc.fdk(g, p, "ramp", prec=32, vol=v0)
    assert rel(vol, ref) <= 1e-4
    out = np.zeros_like(p)
    lib.RampFiltering_GPU(fp(p), fp(out))
    assert lib.ct_compat_last_status() == 0
    rref = orc.ramp_filter(g, p, "ramp", prec=64)
    assert rel(out, rref) <= 1e-5
    q = p.copy()
    lib.RampFiltering_GPU(fp(q), fp(q))
    assert rel(q, rref) <= 1e-5


def test_norm_wrappers(compat):
    """TV_norm_gpu, TV_norm_cpu (V63/host_tool_functions.cu:61,105) and
    L2_norm_cpu (:18) against the oracle's FP64 norms."""
    lib = compat(G)
    x = vol_of(G, seed=7)
    want = orc.tv_norm(x.astype(np.float64), prec=64)
    assert lib.TV_norm_gpu(fp(x)) == pytest.approx(want, rel=1e-10)
    assert lib.TV_norm_cpu(fp(x)) == pytest.approx(want, rel=1e-12)
    a = np.random.default_rng(2).standard_normal(1000).astype(np.float32)
    b = np.random.default_rng(3).standard_normal(1000).astype(np.float32)
    assert lib.L2_norm_cpu(fp(a), fp(b), 1000) == pytest.approx(orc.residual_norm2(a, b, prec=32), rel=1e-12)


def test_multisource_wrappers(compat):
    """Forward_3D_ray_driven_siddon_TBCT and Reconstruction_3D_ray_driven_TBCT
    (TB/voxel/host_IterativeRecon_TBCT_voxel.c:111, :1) on a four-source
    geometry: forward within 1e-5 and one sweep within 1e-4 of the oracle."""
    lib = compat(GMS)
    x = vol_of(GMS, seed=11)
    proj = np.zeros(GMS.proj_shape, dtype=np.float32)
    lib.Forward_3D_ray_driven_siddon_TBCT(fp(x), fp(proj))
    assert lib.ct_compat_last_status() == 0
    ref = orc.forward(GMS, x, prec=32)
    assert rel(proj, ref) <= 1e-5
    vol = np.zeros_like(x)
    lib.Reconstruction_3D_ray_driven_TBCT(fp(vol), fp(proj), ctypes.c_float(0.5))
    assert lib.ct_compat_last_status() == 0
    assert rel(vol, orc.sart_sweep(GMS, np.zeros_like(x), proj, beta=0.5, subset=1, prec=32)) <= 1e-4


def test_save_device_data_to_file(compat, ctx):
    """SaveDeviceDataToFile (V63/host_tool_functions.cu:2) writes raw floats
    from host or device memory."""
    import torch
    lib = compat(G)
    a = np.arange(100, dtype=np.float32)
    d = torch.as_tensor(a).cuda()
    with tempfile.TemporaryDirectory() as td:
        for src, name in ((a, "h.dat"), (None, "d.dat")):
            path = os.path.join(td, name).encode()
            ptr = fp(src) if src is not None else ctypes.cast(ctypes.c_void_p(d.data_ptr()), FP)
            torch.cuda.synchronize()
            lib.SaveDeviceDataToFile(ptr, 100, path)
            assert lib.ct_compat_last_status() == 0
            assert np.array_equal(np.fromfile(path.decode(), dtype=np.float32), a)
