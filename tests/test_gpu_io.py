"""GPU: the file readers and writers and the log transform through the C ABI,
against numpy reading and writing the same files; exact resume from a FISTA
checkpoint; the reference's SaveDeviceDataToFile."""
import ctypes
from dataclasses import replace

import numpy as np
import pytest
import torch

from gpu_amd import core
from gpu_amd.geometry import Geometry
from gpu_amd.phantom import phantom
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

G = Geometry(nx=24, ny=20, nz=16, vx=1.0, ncols=40, nrows=28, du=1.1, nviews=9, dso=90.0, dod=-45.0,
             angle_step_deg=37.3, angle0_deg=-20.0)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def write_views(g, directory, extra_rows, tomo=False, seed=0):
    rng = np.random.default_rng(seed)
    (directory / "Projections").mkdir(parents=True, exist_ok=True)
    data = []
    for j in range(g.nviews):
        a = rng.standard_normal((g.nrows + extra_rows, g.ncols)).astype(np.float32)
        a.tofile(core.projection_path(g, j, str(directory) + "/", tomo=tomo))
        data.append(a)
    return np.stack(data)


@pytest.mark.parametrize("tomo", [False, True])
def test_read_projections_rows_views_and_types(ctx, tmp_path, tomo):
    files = write_views(G, tmp_path, extra_rows=5, tomo=tomo)
    d = str(tmp_path) + "/"
    # all views, rows 3 .. 3+nrows (the reference's fseek offset)
    got = host(ctx.read_projections(G, d, row0=3, tomo=tomo))
    assert np.array_equal(got, files[:, 3:3 + G.nrows])
    # a view range into FP16 / BF16 storage: the FP32 values rounded to nearest even
    for dt in (torch.float16, torch.bfloat16):
        t = ctx.read_projections(G, d, view0=2, nviews=5, row0=0, tomo=tomo, dtype=dt)
        want = torch.from_numpy(np.ascontiguousarray(files[2:7, :G.nrows])).to(dt)
        assert torch.equal(t.cpu(), want)


def test_read_projections_many_chunks(ctx, tmp_path):
    # 64 MiB staging holds 5 views of 2048 x 1600: 23 views take five chunks
    g = Geometry(nx=8, ny=8, nz=8, vx=1.0, ncols=2048, nrows=1600, du=0.1, nviews=23, dso=90.0, dod=-45.0,
                 angle_step_deg=1.0)
    files = write_views(g, tmp_path, extra_rows=0, seed=3)
    got = host(ctx.read_projections(g, str(tmp_path) + "/"))
    assert np.array_equal(got, files)


def test_read_projections_errors(ctx, tmp_path):
    files = write_views(G, tmp_path, extra_rows=0)
    d = str(tmp_path) + "/"
    # asking for rows past the end of the files
    with pytest.raises(core.CtError) as e:
        ctx.read_projections(G, d, row0=1)
    assert e.value.status == 10 and "shorter" in str(e.value)
    missing = core.projection_path(G, 4, d)
    (tmp_path / missing[len(d):]).unlink()
    with pytest.raises(core.CtError) as e:
        ctx.read_projections(G, d)
    assert e.value.status == 10 and "View_" in str(e.value)
    assert files.size > 0


def test_log_transform(ctx):
    rng = np.random.default_rng(1)
    n, h, w = 3, 17, 33
    flat = rng.uniform(900, 1100, (h, w)).astype(np.float32)
    dark = rng.uniform(0, 20, (h, w)).astype(np.float32)
    raw = (rng.uniform(0.0, 1.0, (n, h, w)) * (flat - dark) + dark).astype(np.float32)
    raw[0, 0, :4] = dark[0, :4] - 1.0          # below dark: clamped at the floor
    floor = 1e-6

    def ref(r, f, dk):
        tr = (r.astype(np.float64) - dk) / (f.astype(np.float64) - dk)
        return -np.log(np.maximum(tr, floor))

    got = host(ctx.log_transform(torch.from_numpy(raw).cuda(), torch.from_numpy(flat).cuda(),
                                 torch.from_numpy(dark).cuda(), floor=floor))
    want = ref(raw, flat, dark)
    assert np.allclose(got, want, rtol=2e-6, atol=2e-6)
    assert np.allclose(got[0, 0, :4], -np.log(floor), rtol=1e-6)
    # scalar I0, no dark; FP16 raw in, FP16 out
    i0 = 1e5
    counts = rng.poisson(i0 * np.exp(-rng.uniform(0, 4, (n, h, w)))).astype(np.float32)
    got = host(ctx.log_transform(torch.from_numpy(counts).cuda(), i0=i0))
    want = ref(counts, np.full((h, w), i0, np.float32), np.zeros((h, w), np.float32))
    assert np.allclose(got, want, rtol=2e-6, atol=2e-6)
    c16 = torch.from_numpy(counts / 64).to(torch.float16).cuda()
    out16 = torch.empty_like(c16)
    ctx.log_transform(c16, i0=i0 / 64, out=out16)
    want16 = ref(c16.float().cpu().numpy(), np.full((h, w), i0 / 64, np.float32), np.zeros((h, w), np.float32))
    assert np.allclose(out16.float().cpu().numpy(), want16, rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("dt", [torch.float32, torch.float16, torch.bfloat16, torch.float64])
def test_volume_files_round_trip(ctx, tmp_path, dt):
    x = phantom(G, seed=5, scale=0.3).to(dt).cuda()
    p = tmp_path / "vol.recon"
    ctx.write_volume(G, x, str(p))
    on_disk = np.fromfile(p, dtype=np.float32)
    assert on_disk.size == G.nx * G.ny * G.nz
    assert np.array_equal(on_disk, x.float().cpu().numpy().ravel())
    back = ctx.read_volume(G, str(p), dtype=dt)
    assert torch.equal(back.cpu(), x.float().to(dt).cpu())


def test_large_volume_file_chunks(ctx, tmp_path):
    # 1031 x 48818 voxels: two full 64 MiB chunks of floats and a part chunk
    g = Geometry(nx=1031, ny=1, nz=48818, vx=1.0, ncols=8, nrows=8, du=1.0, nviews=1, dso=1e6, dod=-1.0)
    x = torch.randn(g.vol_shape, device="cuda")
    p = tmp_path / "big.recon"
    ctx.write_volume(g, x, str(p))
    assert np.array_equal(np.fromfile(p, dtype=np.float32), x.cpu().numpy().ravel())
    assert torch.equal(ctx.read_volume(g, str(p)), x)


@pytest.mark.parametrize("dt", [torch.float32, torch.float16])
def test_checkpoint_resume_is_exact(ctx, tmp_path, dt):
    g = G
    b = torch.from_numpy(orc.forward(g, phantom(g, seed=7, scale=0.3).numpy(), prec=32)).cuda()

    def run(x, xp, t, k):
        for _ in range(k):
            t = ctx.sart_fista_iteration(g, x, xp, b, t, beta=0.8, subset=3, fgp_lambda=1e-3, fgp_niter=5)
        return t

    x, xp = torch.zeros(g.vol_shape, device="cuda", dtype=dt), torch.zeros(g.vol_shape, device="cuda", dtype=dt)
    t = run(x, xp, 1.0, 3)
    p = tmp_path / "state.ckpt"
    ctx.save_checkpoint(g, str(p), x, xp, t, 3)
    t_full = run(x, xp, t, 2)
    x2, xp2 = torch.full_like(x, 7.0), torch.full_like(x, -1.0)
    t2, k2 = ctx.load_checkpoint(g, str(p), x2, xp2)
    assert (t2, k2) == (t, 3)
    t2 = run(x2, xp2, t2, 2)
    assert t2 == t_full
    assert torch.equal(x2, x) and torch.equal(xp2, xp)


def test_checkpoint_rejects_bad_files(ctx, tmp_path):
    x = torch.randn(G.vol_shape, device="cuda")
    p = tmp_path / "state.ckpt"
    ctx.save_checkpoint(G, str(p), x, x * 2, 1.5, 4)
    assert not (tmp_path / "state.ckpt.part").exists()
    y, yp = torch.empty_like(x), torch.empty_like(x)
    with pytest.raises(core.CtError) as e:          # another geometry
        ctx.load_checkpoint(replace(G, nviews=10), str(p), y, yp)
    assert e.value.status == 2
    with pytest.raises(core.CtError) as e:          # another storage type
        ctx.load_checkpoint(G, str(p), y.half(), yp.half())
    assert e.value.status == 3
    raw = bytearray(p.read_bytes())
    raw[-5] ^= 0x40                                   # one flipped bit in x_prev
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(bytes(raw))
    with pytest.raises(core.CtError) as e:
        ctx.load_checkpoint(G, str(bad), y, yp)
    assert e.value.status == 10 and "checksum" in str(e.value)
    bad.write_bytes(bytes(raw[:100]))               # truncated
    with pytest.raises(core.CtError) as e:
        ctx.load_checkpoint(G, str(bad), y, yp)
    assert e.value.status == 10


def test_save_device_data_to_file(tmp_path):
    """SaveDeviceDataToFile (V63/host_tool_functions.cu:2) from a device and a host array."""
    lib = core.lib()
    lib.SaveDeviceDataToFile.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
    lib.SaveDeviceDataToFile.restype = None
    lib.ct_compat_last_status.restype = ctypes.c_int
    d = torch.randn(1000, device="cuda")
    p = tmp_path / "dev.dat"
    lib.SaveDeviceDataToFile(ctypes.c_void_p(d.data_ptr()), 1000, str(p).encode())
    assert lib.ct_compat_last_status() == 0
    assert np.array_equal(np.fromfile(p, dtype=np.float32), d.cpu().numpy())
    h = np.arange(77, dtype=np.float32)
    lib.SaveDeviceDataToFile(h.ctypes.data_as(ctypes.c_void_p), 77, str(tmp_path / "host.dat").encode())
    assert lib.ct_compat_last_status() == 0
    assert np.array_equal(np.fromfile(tmp_path / "host.dat", dtype=np.float32), h)
    lib.SaveDeviceDataToFile(h.ctypes.data_as(ctypes.c_void_p), 77, str(tmp_path / "no" / "x.dat").encode())
    assert lib.ct_compat_last_status() == 10
