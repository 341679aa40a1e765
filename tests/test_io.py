"""File names and the objective log (no GPU): the reference's naming of the
per-view projection files (main_recon.cu:199-224), the .recon and
object_func outputs (:438,452,468), and the raw double triples of the
objective file (:467-477), against an independent restatement in numpy
float32 arithmetic and Python %-formatting."""
import numpy as np
import pytest

from gpu_amd import core
from gpu_amd.geometry import Geometry, repo_scale_512

f32 = np.float32


def file_angle(j: int, nviews: int, step: float, a0: float, tomo: bool) -> float:
    a = f32(f32(j) * f32(step)) + f32(a0)
    if tomo and j >= nviews // 2:
        a = f32(f32(f32(180.0) + f32(f32(j - nviews // 2) * f32(step))) + f32(a0))
    if a < 0:
        a = f32(a + f32(360.0))
    return float(a)


GEOMS = [
    repo_scale_512(),
    Geometry(nx=8, ny=8, nz=8, vx=1.0, ncols=16, nrows=8, du=1.0, nviews=40, dso=100.0, dod=-50.0,
             angle_step_deg=1.7, angle0_deg=-33.3),
    Geometry(nx=8, ny=8, nz=8, vx=0.125, ncols=16, nrows=8, du=1.0, nviews=7, dso=100.0, dod=-50.0,
             angle_step_deg=0.1, angle0_deg=12.345),
]


@pytest.mark.parametrize("gi", range(len(GEOMS)))
@pytest.mark.parametrize("tomo", [False, True])
def test_projection_file_names(gi, tomo):
    g = GEOMS[gi]
    step = g.angle_step_deg if g.angle_step_deg is not None else 360.0 / g.nviews
    for j in sorted({0, 1, g.nviews // 2 - 1, g.nviews // 2, g.nviews - 1}):
        want = "/data/" + "Projections/View_%.02f.dat" % file_angle(j, g.nviews, step, g.angle0_deg, tomo)
        assert core.projection_path(g, j, "/data/", tomo=tomo) == want


def test_negative_angles_wrap():
    g = GEOMS[1]
    # view 0 sits at -33.3 deg, written as 326.70
    assert core.projection_path(g, 0, "") == "Projections/View_326.70.dat"
    assert core.projection_path(g, 39, "") == "Projections/View_%.02f.dat" % file_angle(39, 40, 1.7, -33.3, False)


@pytest.mark.parametrize("gi", range(len(GEOMS)))
def test_output_names(gi):
    g = GEOMS[gi]
    step = g.angle_step_deg if g.angle_step_deg is not None else 360.0 / g.nviews
    a0 = float(f32(g.angle0_deg))
    a1 = float(f32(a0) + f32(f32(g.nviews - 1) * f32(step)))
    um = float(f32(g.vx) * f32(1e3))
    lam, l1, lip = 1e-3, 0.25, 32.0
    tail = "_TV_%.2f_L1_%.2f_Lip_%.2f.recon" % (f32(lam), f32(l1), f32(lip))
    head = "/home/u/%s/%s_%d_%d_%d_%.0fum_" % ("/Recon_Test/", "Ellipsoid", g.nx, g.ny, g.nz, um)
    assert core.recon_path(g, "/home/u/", "/Recon_Test/", "Ellipsoid", -1, lam, l1, lip) == \
        head + "new_view_%d_(%.0f,%.0f)" % (g.nviews, a0, a1) + tail
    assert core.recon_path(g, "/home/u/", "/Recon_Test/", "Ellipsoid", 6, lam, l1, lip) == \
        head + "iterative_6_view_%d_(%.0f,%.0f)" % (g.nviews, a0, a1) + tail
    assert core.objective_path(g, "/home/u/", "/Recon_Test/", "Ellipsoid", lam, lip) == \
        "/home/u//Recon_Test//object_func_Ellipsoid_view_%d_(%.0f,%.0f)_TV_%.2f_Lip_%.2f.bin" % (
            g.nviews, a0, a1, f32(lam), f32(lip))


def test_reference_units_give_reference_name():
    # the reference's constants: 100 um voxels given in metres, 1024^2 x 256
    g = Geometry(nx=1024, ny=1024, nz=256, vx=1.0e-4, ncols=2048, nrows=512, du=1.2e-4, nviews=1, dso=1.0,
                 dod=-1.0, angle_step_deg=1.0, angle0_deg=0.0)
    name = core.recon_path(g, "/home/huifeng/CUDA_multiGPU/CBCT/", "/Recon_Test/", "Ellipsoid", -1, 0.0, 0.0,
                           32.0, unit_um=1e6)
    assert name == ("/home/huifeng/CUDA_multiGPU/CBCT//Recon_Test//Ellipsoid_1024_1024_256_100um_new_view_1_(0,0)"
                    "_TV_0.00_L1_0.00_Lip_32.00.recon")


def test_path_buffer_too_small_and_bad_view():
    import ctypes
    g = GEOMS[1]
    gc = g.to_c()
    buf = ctypes.create_string_buffer(8)
    assert core.lib().ct_projection_path(ctypes.byref(gc), 0, 0, b"/data/", buf, 8) == 1
    buf = ctypes.create_string_buffer(256)
    assert core.lib().ct_projection_path(ctypes.byref(gc), g.nviews, 0, b"/data/", buf, 256) == 1


def test_objective_file_is_raw_doubles(tmp_path):
    vals = [3.5, 1e-300, -2.25, 1.0 / 3.0, 7e10, 0.0]
    p = tmp_path / "object_func.bin"
    core.write_objective(str(p), vals)
    got = np.fromfile(p, dtype=np.float64)
    assert got.tolist() == vals
    with pytest.raises(core.CtError) as e:
        core.write_objective(str(tmp_path / "missing_dir" / "x.bin"), vals)
    assert e.value.status == 10
