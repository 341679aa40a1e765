#!/usr/bin/env python3
"""Benchmark: fwd+back GRays/s and seconds per SART-FISTA-TV iteration at
512^3 voxels / 1024 views of 1024^2, FP32 storage vs FP16 storage.

Workload (synthetic, seeded): 512^3 volume of 0.5 mm voxels holding the
10-ellipsoid phantom scaled to +-100 mm; 1024 views over 360 degrees onto a
1024x1024 detector of 0.4 mm pixels, SID 1000 mm / SDD 1500 mm; Poisson noise
with I0 = 1e5 photons per ray, then the log transform.  One step = one outer
iteration: a SART sweep over all 1024 views in the reference's view-by-view
order (relaxation 2/Lip with Lip from 30 power iterations of the operator one
sweep applies, in place of the reference's constant Lip_con = 32), TV-FGP with
20 inner iterations (lambda 1e-3) and the FISTA momentum, all device
resident.  Inputs are in HBM before timing starts.

value = fwd+back GRays/s of the FP32 path = rays per sweep / (time of the
Siddon correction kernel + time of the voxel-driven update kernel), both timed
live with HIP events on the context stream.  The same run reports the FP16
and BF16 paths, seconds per iteration of each, the roofline of the dominant
kernel and of both sweep kernels in FP32 and FP16, the error of three
iterations against the FP64 path, FDK (Hann) over the same projections, the
autograd pair (ray sums + matched adjoint) and the multi-source pair on
samples of views, other beam styles, file I/O, the 128^3 / 180-view SART
workload, and the CPU oracle (OpenMP restatement of the reference kernels) on
bounded samples of the same workloads.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "fwd+back GRays/s; s per SART-FISTA iter at 512³/1024 views; FP16 vs FP32"
HBM_PEAK_GBS = 8000.0
VALU_ISSUE_PER_S = 256 * 2.4e9     # wave64 VALU instructions per second (1 per CU per clock)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_path(ctx, g, meas, dtype, steps, warmup, beta, lam, niter):
    import torch
    x = torch.zeros(g.vol_shape, device="cuda", dtype=dtype)
    xp = torch.zeros_like(x)
    t = 1.0
    for _ in range(warmup):
        t = ctx.sart_fista_iteration(g, x, xp, meas, t, beta=beta, subset=1, fgp_lambda=lam, fgp_niter=niter)
    torch.cuda.synchronize()
    ctx.profile(True, reset=True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        t = ctx.sart_fista_iteration(g, x, xp, meas, t, beta=beta, subset=1, fgp_lambda=lam, fgp_niter=niter)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    prof = ctx.profile_read()
    ctx.profile(False, reset=False)
    return ms / steps, prof, xp


def accuracy(ctx, g, x_true, meas32, iters, beta, lam, niter):
    """Per-iteration RMSE / relative L2 of the FP32, FP16 and BF16 paths against
    the FP64 path (same data, same schedule, from zero) and against the
    phantom.  Returns the record and the seconds per iteration of each path."""
    import torch
    paths = {"f64": torch.float64, "f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}
    state = {}
    for k, dt in paths.items():
        meas = meas32 if dt in (torch.float32, torch.float64) else meas32.to(dt)
        state[k] = [torch.zeros(g.vol_shape, device="cuda", dtype=dt), torch.zeros(g.vol_shape, device="cuda", dtype=dt),
                    1.0, meas, 0.0]
    out = {"workload": f"512^3 / 1024 views SART(subset 1, beta {beta:.6f}) + TV-FGP({niter}, {lam}) + FISTA, "
                       "from zero",
           "reference": "f64 path on the GPU (FP64 instantiation of the same kernels)", "iterations": []}
    truth = x_true.double()
    tn = float(torch.linalg.vector_norm(truth))
    for it in range(iters):
        rec = {"iteration": it + 1}
        for k in paths:
            x, xp, t, meas, _ = state[k]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            t = ctx.sart_fista_iteration(g, x, xp, meas, t, beta=beta, subset=1, fgp_lambda=lam, fgp_niter=niter)
            e1.record()
            torch.cuda.synchronize()
            state[k][2] = t
            state[k][4] += e0.elapsed_time(e1) / 1e3
            log(f"accuracy: iteration {it + 1} {k} {e0.elapsed_time(e1) / 1e3:.2f} s")
        ref = state["f64"][1]          # x_k lives in x_prev after the momentum step
        rn = float(torch.linalg.vector_norm(ref))
        for k in paths:
            xk = state[k][1].double()
            d = xk - ref
            rec[k] = {"rmse_vs_f64": float(torch.sqrt((d * d).mean())),
                      "rel_l2_vs_f64": float(torch.linalg.vector_norm(d)) / rn,
                      "rmse_vs_phantom": float(torch.sqrt(((xk - truth) ** 2).mean())),
                      "rel_l2_vs_phantom": float(torch.linalg.vector_norm(xk - truth)) / tn}
        out["iterations"].append(rec)
    out["seconds_per_iteration"] = {k: state[k][4] / iters for k in paths}
    del state
    torch.cuda.empty_cache()
    return out


def adam_multisource(ctx, iters: int, storage: str, ref_iters: int = 0) -> dict:
    """The multi-source gradient workload at its own size: 256^3 at 0.75 mm,
    48 point sources 3 mm apart on the axis, 720 views of 1024 x 128 at
    0.5 mm (34560 images, 4.5e9 rays); Adam (lr 1e-3) on
    0.5 ||A x - b||^2 + 1e-3 TV_eps(x), eps 1e-8, through the autograd
    projector (forward = ray sums, backward = the matched adjoint, the rays
    re-walked instead of storing per-ray intermediates), volume stored in
    `storage` with FP32 master weights, from zero.  b = A x_true of the seeded
    phantom (FP32).  Reports seconds per iteration (HIP events), the loss per
    iteration and the distance to the phantom.  With `ref_iters`, the same
    recurrence first runs that many iterations entirely in FP64 (the GPU's
    FP64 instantiation of the projector pair, FP64 master weights and
    measurements) and the run is compared with it after as many iterations:
    an FP64 CPU reconstruction at this size would take days, and the 24^3
    test compares with the FP64 CPU recurrence."""
    import torch
    from gpu_amd.autograd import project, tv_eps
    from gpu_amd.geometry import multisource_256
    from gpu_amd.phantom import phantom
    g = multisource_256()
    sdt = {"bf16": torch.bfloat16, "f16": torch.float16, "f32": torch.float32}[storage]
    x_true = phantom(g, seed=0, scale=2.0, device="cuda")
    b = ctx.forward(g, x_true)
    lam, eps = 1e-3, 1e-8

    def run(n, store, master, meas, tag, snap_at=0):
        x = torch.zeros(g.vol_shape, device="cuda", dtype=master, requires_grad=True)
        opt = torch.optim.Adam([x], lr=1e-3)
        losses, times, snap = [], [], None
        for k in range(n):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            opt.zero_grad()
            xs = x.to(store)
            r = project(xs, g, ctx, out_dtype=meas.dtype) - meas
            loss = 0.5 * torch.linalg.vector_norm(r, dtype=torch.float64) ** 2 + lam * tv_eps(xs, eps, ctx)
            loss.backward()
            opt.step()
            e1.record()
            torch.cuda.synchronize()
            losses.append(float(loss))
            times.append(e0.elapsed_time(e1) / 1e3)
            log(f"adam {tag}: iteration {k + 1} loss {losses[-1]:.6e} {times[-1]:.2f} s")
            del r, loss, xs
            if k + 1 == snap_at:
                snap = x.detach().double().clone()
        return x.detach(), losses, times, snap

    ref = None
    if ref_iters > 0:
        b64 = b.double()
        x64, loss64, t64, _ = run(ref_iters, torch.float64, torch.float64, b64, "f64")
        del b64
        torch.cuda.empty_cache()
        ref = (x64.double(), loss64, t64)
    x, losses, times, snap = run(iters, sdt, torch.float32, b, storage, snap_at=ref_iters)
    xt = x_true.double()
    d = x.double() - xt
    out = {"workload": "multi-source Adam, 256^3 x 0.75 mm, 48 sources x 720 views of 1024 x 128, lr 1e-3, "
                       "0.5||Ax-b||^2 + 1e-3 TV_eps",
           "storage": storage, "iterations": iters,
           "s_per_iteration": sum(times[1:]) / max(len(times) - 1, 1), "first_iteration_s": times[0],
           "loss": losses, "rmse_vs_phantom": float(torch.sqrt((d * d).mean())),
           "rel_l2_vs_phantom": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(xt))}
    if ref is not None and snap is not None:
        x64, loss64, t64 = ref
        dd = snap - x64
        out["vs_f64"] = {
            "iterations": ref_iters, "rmse": float(torch.sqrt((dd * dd).mean())),
            "rel_l2": float(torch.linalg.vector_norm(dd) / torch.linalg.vector_norm(x64)),
            "max_rel_loss_diff": max(abs(a - c) / abs(c) for a, c in zip(losses[:ref_iters], loss64)),
            "f64_s_per_iteration": sum(t64) / len(t64),
            "reference": "the same Adam recurrence in FP64 on the GPU (FP64 projector pair, master weights and "
                         "measurements)"}
    return out


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, meas_host, beta, lam, niter, sweep_views, full=False):
    """The oracle (OpenMP restatement of the reference kernels, FP32, -O3, IEEE)
    on the host cores, on samples of the same workloads:

    * the 512^3 / 1024-view SART sweep's first `sweep_views` views (correction +
      voxel update), extrapolated to the whole sweep and iteration (labelled);
    * TV-FGP inner iterations at 512^3 and one FISTA momentum step;
    * FDK (Hann) of 4 views at 512^3;
    * the 128^3 / 180-view SART workload: one SART iteration (subsets of 20)
      by default, the full ten iterations three times (median) with `full`.
    """
    import numpy as np
    from gpu_amd.geometry import plain_sart_128
    from gpu_amd.phantom import phantom
    from oracle import oracle as orc
    threads = orc.num_threads()
    out = {"unit": "GRays/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
           "nproc": os.cpu_count()}
    # SART pair on the first views of the sweep
    v = np.zeros(g.vol_shape, dtype=np.float32)
    t0 = time.perf_counter()
    for j in range(sweep_views):
        c = orc.sart_correction(g, v, meas_host[j:j + 1], view0=j, nviews=1, prec=32)
        v = orc.backproject_voxel(g, v, c, view0=j, nviews=1, beta=beta, mode=1, prec=32)
    dt = time.perf_counter() - t0
    rays = sweep_views * g.nrows * g.ncols
    out["value"] = rays / dt / 1e9
    out["sample"] = (f"first {sweep_views} views of the 512^3 / 1024-view SART sweep (correction + voxel update), "
                     f"FP32, {dt:.1f} s, OpenMP {threads} threads")
    sweep_s = dt / sweep_views * g.nviews
    # TV-FGP and FISTA at 512^3 (sampled: 2 and 4 inner iterations, the
    # difference giving the cost per inner iteration apart from the set-up)
    if full:
        t0 = time.perf_counter()
        orc.tv_fgp(v, lam, niter, prec=32)
        fgp_s = time.perf_counter() - t0
    else:
        t0 = time.perf_counter()
        orc.tv_fgp(v, lam, 2, prec=32)
        t2 = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.tv_fgp(v, lam, 4, prec=32)
        t4 = time.perf_counter() - t0
        fgp_s = t2 + (t4 - t2) / 2 * (niter - 2)
    xp = np.zeros_like(v)
    t0 = time.perf_counter()
    orc.fista_momentum(v, xp, 0.5, prec=32)
    fista_s = time.perf_counter() - t0
    out["sart_fista_tv_512"] = {
        "sweep_s_extrapolated": round(sweep_s, 1), "tv_fgp_s": round(fgp_s, 2), "fista_s": round(fista_s, 3),
        "s_per_iter_extrapolated": round(sweep_s + fgp_s + fista_s, 1),
        "note": f"sweep extrapolated x{g.nviews / sweep_views:.0f} from {sweep_views} views"
                + ("" if full else f"; TV-FGP({niter}) extrapolated from runs of 2 and 4 inner iterations")}
    # FDK of 4 views
    g4 = g.with_views(4)
    t0 = time.perf_counter()
    orc.fdk(g4, meas_host[:4], "hann", prec=32)
    out["fdk_512_4_views_s"] = round(time.perf_counter() - t0, 2)
    # the 128^3 / 180-view SART workload
    g1 = plain_sart_128()
    x1 = phantom(g1, seed=0).numpy()
    b1 = orc.forward(g1, x1, prec=32)
    iters, runs = (10, 3) if full else (1, 1)
    times = []
    for _ in range(runs):
        r = np.zeros_like(x1)
        t0 = time.perf_counter()
        for _ in range(iters):
            r = orc.sart_sweep(g1, r, b1, beta=1.0, subset=20, order="interleaved", prec=32)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    out["plain_sart_128"] = {"iterations": iters, "runs": runs, "median_s": round(med, 2),
                             "s_per_iter": round(med / iters, 2),
                             "grays": round(g1.rays * iters / med / 1e9, 5)}
    return out


def plain_sart_128_gpu(ctx) -> dict:
    """The CPU-runnable workload on the GPU: 128^3 at 1 mm, 180 views of 256^2,
    ten SART iterations (relaxation 1, interleaved subsets of 20 views) from
    zero on the phantom's clean projections, FP32 (the CPU baseline runs the
    same)."""
    import torch
    from gpu_amd.geometry import plain_sart_128
    from gpu_amd.phantom import phantom
    g1 = plain_sart_128()
    x1 = phantom(g1, seed=0, device="cuda")
    b1 = ctx.forward(g1, x1)
    v = torch.zeros(g1.vol_shape, device="cuda")
    ctx.sart_sweep(g1, v, b1, beta=1.0, subset=20, order="interleaved")   # warm-up
    v.zero_()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(10):
        ctx.sart_sweep(g1, v, b1, beta=1.0, subset=20, order="interleaved")
    e1.record()
    torch.cuda.synchronize()
    s = e0.elapsed_time(e1) / 1e3
    return {"iterations": 10, "s": round(s, 4), "s_per_iter": round(s / 10, 5),
            "grays": round(g1.rays * 10 / s / 1e9, 4)}


def style_figures(ctx, g, x32) -> dict:
    """SART pair (correction + voxel-driven update) per view for the other
    beam styles on the same volume: multi-slice fan beam and parallel beam,
    with one detector row per slice (512 rows of 0.5 mm), 8 views."""
    from dataclasses import replace

    import torch

    out = {}
    nv = 8
    for name, style in (("fan", 1), ("parallel", 2)):
        gs = replace(g, style=style, nrows=g.nz, dv=g.vz or g.vx).with_views(nv)
        meas = ctx.forward(gs, x32)
        vol = torch.zeros_like(x32)
        ctx.sart_sweep(gs, vol, meas, beta=0.0625)          # warm-up
        ctx.profile(True, reset=True)
        ctx.sart_sweep(gs, vol, meas, beta=0.0625)
        torch.cuda.synchronize()
        pf = ctx.profile_read()
        ctx.profile(False)
        fwd, bp = pf["forward"][0] / nv, pf["voxel_bp"][0] / nv
        out[name] = {"grays": round(gs.nrows * gs.ncols / ((fwd + bp) * 1e-3) / 1e9, 4),
                     "correction_ms": round(fwd, 4), "update_ms": round(bp, 4)}
    out["note"] = "512^3, 8 views of 1024 x 512 (rows at the slice centres), view-by-view SART sweep"
    return out


def multisource_figures(ctx) -> dict:
    """The multi-source autograd workload at its own size (48 point sources 3
    mm apart on the axis, 720 views of 1024 x 128 at 0.5 mm, 256^3 at 0.75 mm,
    BF16 storage, FP32 accumulation): ray sums then the matched adjoint on 4
    batches of 2 views (96 images each) spread over the orbit, and the time
    one Adam iteration's forward + gradient projection would take at that
    rate."""
    import torch
    from gpu_amd.geometry import multisource_256
    from gpu_amd.phantom import phantom

    gm = multisource_256()
    # the seeded phantom scaled to +-80 mm inside the +-96 mm volume
    x = phantom(gm, seed=0, scale=2.0, device="cuda").to(torch.bfloat16)
    nbatch, bviews = 4, 2
    # warm-up (first launches of these kernel instantiations)
    ctx.adjoint(gm, ctx.forward(gm, x, view0=0, nviews=1), view0=0, nviews=1)
    ctx.profile(True, reset=True)
    for j in range(nbatch):
        v0 = j * (gm.nviews // nbatch)
        pj = ctx.forward(gm, x, view0=v0, nviews=bviews)
        ctx.adjoint(gm, pj, view0=v0, nviews=bviews)
    torch.cuda.synchronize()
    pa = ctx.profile_read()
    ctx.profile(False)
    nimg = nbatch * bviews * gm.nsources
    f_ms, a_ms = pa["forward"][0] / nimg, pa["adjoint"][0] / nimg
    rays = gm.nrows * gm.ncols
    return {"grays": round(rays / ((f_ms + a_ms) * 1e-3) / 1e9, 4),
            "forward_ms_per_image": round(f_ms, 5), "adjoint_ms_per_image": round(a_ms, 5),
            "est_s_per_adam_iteration": round((f_ms + a_ms) * gm.nviews * gm.nsources / 1e3, 2),
            "sample": f"{nbatch} batches of {bviews} views x {gm.nsources} sources, 256^3 BF16, 720 views of "
                      "1024 x 128; one image = one (view, source) pair"}


def io_figures(ctx, g, meas32, x32) -> dict:
    """The callers either side of the path, on the bench's own data: 64 view
    files read back into HBM (from the page cache: the files were just
    written), the log transform of the whole 1024-view stack of counts, and
    the volume and FISTA-checkpoint files written and read."""
    import tempfile

    import torch
    from gpu_amd import core

    out = {}
    nio = 64
    gio = g.with_views(nio)
    host = meas32[:nio].cpu().numpy()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    with tempfile.TemporaryDirectory() as td:
        d = td + "/"
        os.makedirs(d + "Projections")
        for j in range(nio):
            host[j].tofile(core.projection_path(gio, j, d))
        buf = torch.empty((nio, g.nrows, g.ncols), device="cuda")
        ctx.read_projections(gio, d, buf)                    # first pass warms the page cache
        t0 = time.perf_counter()
        ctx.read_projections(gio, d, buf)
        t_read = time.perf_counter() - t0
        assert torch.equal(buf, meas32[:nio])
        out["read_views_GBps"] = round(buf.numel() * 4 / t_read / 1e9, 2)
        del buf
        counts = torch.exp(-meas32) * 1e5
        lines = torch.empty_like(counts)
        ctx.log_transform(counts, i0=1e5, out=lines)
        torch.cuda.synchronize()
        ev[0].record()
        ctx.log_transform(counts, i0=1e5, out=lines)
        ev[1].record()
        torch.cuda.synchronize()
        t_log = ev[0].elapsed_time(ev[1]) * 1e-3
        gbps = counts.numel() * 8 / t_log / 1e9
        out["log_transform"] = {"ms": round(t_log * 1e3, 3), "GBps": round(gbps, 1),
                                "hbm_frac": round(gbps / HBM_PEAK_GBS, 3)}
        del counts, lines
        vol_path = d + "vol.recon"
        t0 = time.perf_counter()
        ctx.write_volume(g, x32, vol_path)
        t_w = time.perf_counter() - t0
        t0 = time.perf_counter()
        back = ctx.read_volume(g, vol_path)
        t_r = time.perf_counter() - t0
        assert torch.equal(back, x32)
        out["volume_write_GBps"] = round(x32.numel() * 4 / t_w / 1e9, 2)
        out["volume_read_GBps"] = round(x32.numel() * 4 / t_r / 1e9, 2)
        ck = d + "state.ckpt"
        t0 = time.perf_counter()
        ctx.save_checkpoint(g, ck, x32, back, 1.5, 3)
        t_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        ctx.load_checkpoint(g, ck, back, back.clone())
        t_l = time.perf_counter() - t0
        out["checkpoint_save_s"] = round(t_s, 3)
        out["checkpoint_load_s"] = round(t_l, 3)
    out["note"] = ("views: 64 raw float32 files of 1024^2 read through pinned buffers from the page cache; "
                   "log transform: 1024 x 1024^2 counts -> line integrals, bytes = 8 per pixel; volume: 512^3 "
                   "float32 .recon; checkpoint: x and x_prev with header and checksum")
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16-storage path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-views", type=int, default=6, help="SART views of the CPU baseline's sweep sample")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU baseline over the whole plan: the 128^3 workload's ten iterations three times "
                         "(median), TV-FGP(20) at 512^3, 32 sweep views")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 reference iterations (accuracy)")
    ap.add_argument("--acc-iters", type=int, default=3,
                    help="SART-FISTA-TV iterations of the accuracy pass (FP64, FP32, FP16, BF16 from zero)")
    ap.add_argument("--accuracy", type=int, default=0,
                    help="accuracy mode: N SART-FISTA-TV iterations in FP64 (the GPU FP64 instantiation as the "
                         "oracle), FP32, FP16 and BF16; prints per-iteration RMSE / relative L2 against FP64 and "
                         "the phantom as one JSON line, then exits")
    ap.add_argument("--adam", type=int, default=0,
                    help="multi-source gradient mode: N Adam iterations of the 256^3 / 48-source / 720-view "
                         "workload (see --adam-storage); prints one JSON line, then exits")
    ap.add_argument("--adam-storage", choices=("bf16", "f16", "f32"), default="bf16")
    ap.add_argument("--adam-ref", type=int, default=0,
                    help="with --adam: first run this many iterations of the same recurrence in FP64 and report "
                         "the run's distance to it after as many iterations")
    ap.add_argument("--beta", type=float, default=0.0,
                    help="SART relaxation; 0 (default) takes 2/Lip from 30 power iterations of the operator one "
                         "view-by-view sweep applies")
    ap.add_argument("--pmc-sample", type=int, default=0,
                    help="counter-collection mode: SART sweeps over N views from 0 and from 90 degrees, an FDK over "
                         "the first N views and two TV-FGP inner iterations at full shape, plus a 1 GiB device "
                         "copy (calibration); exits without the JSON line")
    ap.add_argument("--pmc-dtype", choices=("f32", "f16"), default="f32",
                    help="storage type of the counter-collection sweep")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from gpu_amd.core import Context
    from gpu_amd.geometry import repo_scale_512
    from gpu_amd.phantom import phantom, poisson_log

    g = repo_scale_512()
    # FGP argument lambda = 1e-3, 20 inner iterations; the SART-FISTA step
    # beta = 2/Lip with Lip from the power iteration of the operator that one
    # view-by-view sweep applies (it replaces the reference's constant
    # Lip_con = 32, V63/main_recon.cu:58,161)
    lam, niter = 1e-3, 20
    beta = args.beta if args.beta > 0 else 0.0625
    ctx = Context(0)
    if args.adam:
        print(json.dumps(adam_multisource(ctx, args.adam, args.adam_storage, args.adam_ref)), flush=True)
        return 0
    if args.pmc_sample:
        # counter-collection mode: a cheap input (one ellipsoid-free ball) so the
        # profiled run has few dispatches besides the kernels of interest
        zz, yy, xx = torch.meshgrid(*(torch.arange(n, device="cuda", dtype=torch.float32) - n / 2 + 0.5
                                      for n in g.vol_shape), indexing="ij")
        sdt = torch.float16 if args.pmc_dtype == "f16" else torch.float32
        ball = (((xx * xx + yy * yy + zz * zz) < (0.4 * g.nx) ** 2).to(torch.float32) * 0.02).to(sdt)
        gs = g.with_views(args.pmc_sample)      # the first N views of the orbit
        meas = ctx.forward(gs, ball)
        # calibration: the library's own element copy (k_convert F32 -> F32) of
        # 2^28 floats, 1 GiB read + 1 GiB written
        src = torch.ones(1 << 28, device="cuda")
        dst = torch.empty_like(src)
        ctx.convert(src, dst)
        torch.cuda.synchronize()
        vol = torch.zeros(g.vol_shape, device="cuda", dtype=sdt)
        ctx.sart_sweep(gs, vol, meas, beta=beta, subset=1)      # the sweep's kernels, as timed
        # the same number of views from 90 degrees on: rays along y, which the
        # sweep walks by column on its transposed z-quad copy
        from dataclasses import replace
        g90 = replace(gs, angle0_deg=90.0)
        ctx.sart_sweep(g90, vol, ctx.forward(g90, ball), beta=beta, subset=1)
        # FDK's kernels (weighted filter + backprojection) over the same views
        vol32 = torch.zeros(g.vol_shape, device="cuda")
        ctx.fdk(gs, meas.float(), vol32, window="hann")
        # two TV-FGP inner iterations (one fused step each) on that volume
        ctx.tv_fgp(vol32, 1e-3, 2)
        torch.cuda.synchronize()
        log(f"pmc sample done: {args.pmc_sample} views")
        return 0
    log("setup: phantom + projections + noise")
    x_true = phantom(g, seed=0, scale=2.5, device="cuda")
    clean = ctx.forward(g, x_true)
    meas32 = poisson_log(clean, i0=1e5, seed=1)
    del clean
    torch.cuda.synchronize()

    lip = None
    if args.beta <= 0:
        log("SART step: 30 power iterations of the view-by-view sweep operator")
        t0 = time.perf_counter()
        lip = ctx.sart_lipschitz(g, torch.float32, niter=30, seed=1, subset=1)
        beta = 2.0 / lip
        lip_s = time.perf_counter() - t0
        log(f"SART step: Lip {lip:.6f}, beta {beta:.6f} ({lip_s:.1f} s)")
    if args.accuracy:
        print(json.dumps(accuracy(ctx, g, x_true, meas32, args.accuracy, beta, lam, niter)), flush=True)
        return 0

    log(f"FP32 path: warmup {args.warmup}, steps {args.steps}")
    ms32, prof32, x32 = run_path(ctx, g, meas32, torch.float32, args.steps, args.warmup, beta, lam, niter)
    fwd_ms, fwd_n = prof32["forward"]
    bp_ms, bp_n = prof32["voxel_bp"]
    rays_per_sweep = g.rays
    grays = rays_per_sweep * args.steps / ((fwd_ms + bp_ms) * 1e-3) / 1e9
    log(f"FP32: {ms32 / 1e3:.3f} s/iter, forward {fwd_ms / fwd_n:.3f} ms x {fwd_n}, "
        f"voxel bp {bp_ms / bp_n:.3f} ms x {bp_n}, GRays/s {grays:.3f}")

    V, P = g.nvox, g.nrows * g.ncols
    # HBM bytes and VALU instructions per launch of the sweep kernels, from
    # the committed rocprofv3 counter passes (they cannot be read from inside
    # this process); the live time turns the VALU count into an issue-rate
    # fraction
    tf = ROOT / "profiles" / "traffic.json"
    tj = json.loads(tf.read_text()) if tf.exists() else {}

    def roofline(name, avg_ms, abytes, key):
        """Roofline of one sweep kernel class (one launch = one view)."""
        achieved = abytes / (avg_ms * 1e-3) / 1e9
        vi = tj.get("valu_insts_per_launch", {}).get(key)
        return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": tj.get("hbm_bytes_per_launch", {}).get(key),
                "algorithmic_bytes_per_launch": abytes, "ms_per_launch": round(avg_ms, 4),
                "valu_issue_frac": round(vi / (VALU_ISSUE_PER_S * avg_ms * 1e-3), 4) if vi else None}

    # algorithmic bytes per view: the correction reads f and b and writes c;
    # the update reads c and reads and writes f
    rl32 = {"forward": roofline("forward", fwd_ms / fwd_n, 4 * V + 4 * P + 4 * P, "forward"),
            "voxel_bp": roofline("voxel_bp", bp_ms / bp_n, 4 * P + 2 * 4 * V, "voxel_bp")}
    dom = max(rl32, key=lambda k: rl32[k]["ms_per_launch"])

    result = {
        "metric": METRIC,
        "value": round(grays, 4),
        "unit": "GRays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 10-ellipsoid phantom, Poisson I0=1e5 + log",
        "config": {
            "workload": "SART-FISTA-TV iteration, 512^3 x 0.5 mm, 1024 views x 1024^2 x 0.4 mm, SID/SDD 1000/1500 mm",
            "sart": (f"view by view (subset 1), relaxation {beta:.6e} = 2/Lip, Lip = {lip:.6f} from 30 power "
                     "iterations of the sweep operator sum_v D_v^-1 B_v^T M_v^-1 A_v (FP32 projections, FP64 "
                     "norms)") if lip else
                    f"view by view (subset 1), relaxation {beta}",
            "tv_fgp": {"inner_iterations": niter, "lambda": lam},
            "rays_per_sweep": rays_per_sweep,
        },
        "s_per_iter": {"f32": round(ms32 / 1e3, 4)},
        "kernels_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof32.items() if v[1]},
        "roofline": rl32[dom],
        "roofline_sweep_kernels": {"f32": rl32},
    }

    # secondary figures on the same inputs: FDK (Hann) over all views, and the
    # autograd pair (ray sums + matched adjoint) on a sample of views.  One
    # untimed FDK first: the kernels' first launch loads their code and the
    # filtered-view buffer is allocated once per context.
    ctx.fdk(g, meas32, window="hann")
    ctx.profile(True, reset=True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    vol_fdk = ctx.fdk(g, meas32, window="hann")
    e1.record()
    torch.cuda.synchronize()
    fdk_s = e0.elapsed_time(e1) / 1e3
    pf = ctx.profile_read()
    result["fdk"] = {"s": round(fdk_s, 4), "window": "hann", "views": g.nviews,
                     "kernels_ms": {k: round(v[0], 3) for k, v in pf.items() if v[1]}}
    log(f"FDK (Hann, {g.nviews} views): {fdk_s:.3f} s")
    del vol_fdk
    nbatch, bviews = 4, 8
    nsamp = nbatch * bviews
    ctx.profile(True, reset=True)
    for j in range(nbatch):
        v0 = j * (g.nviews // nbatch)
        pj = ctx.forward(g, x32, view0=v0, nviews=bviews)
        ctx.adjoint(g, pj, view0=v0, nviews=bviews)
    torch.cuda.synchronize()
    pa = ctx.profile_read()
    t_pair = (pa["forward"][0] + pa["adjoint"][0]) / nsamp   # ms per view
    result["autograd_pair"] = {
        "grays": round(g.nrows * g.ncols / (t_pair * 1e-3) / 1e9, 4),
        "forward_ms": round(pa["forward"][0] / nsamp, 4), "adjoint_ms": round(pa["adjoint"][0] / nsamp, 4),
        "sample": f"{nbatch} batches of {bviews} consecutive views spread over the orbit: ray sums, then the "
                  "matched adjoint of the batch"}
    log(f"autograd pair: forward {pa['forward'][0] / nsamp:.3f} ms + adjoint {pa['adjoint'][0] / nsamp:.3f} ms per view")
    ctx.profile(False)
    result["plain_sart_128"] = plain_sart_128_gpu(ctx)
    log(f"128^3 SART: {result['plain_sart_128']['s_per_iter']} s per iteration")
    result["beam_styles"] = style_figures(ctx, g, x32)
    log("beam styles: " + ", ".join(f"{k} {v['grays']} GRays/s" for k, v in result["beam_styles"].items()
                                     if isinstance(v, dict)))
    result["multisource_pair"] = multisource_figures(ctx)
    log(f"multi-source pair: {result['multisource_pair']['grays']} GRays/s")
    result["io"] = io_figures(ctx, g, meas32, x32)
    log("files: " + ", ".join(f"{k} {v}" for k, v in result["io"].items() if k != "note"))

    if not args.no_fp16:
        # the reduced-precision paths: FP16 and BF16 storage, FP32 accumulation,
        # the same schedule and step
        for tag, dt in (("f16", torch.float16), ("bf16", torch.bfloat16)):
            log(f"{tag.upper()} path")
            meas_r = meas32.to(dt)
            ms_r, prof_r, x_r = run_path(ctx, g, meas_r, dt, args.steps, args.warmup, beta, lam, niter)
            f_r = prof_r["forward"][0] + prof_r["voxel_bp"][0]
            grays_r = rays_per_sweep * args.steps / (f_r * 1e-3) / 1e9
            diff = (x_r.float() - x32).double()
            rmse = float(torch.sqrt((diff * diff).mean()))
            result["s_per_iter"][tag] = round(ms_r / 1e3, 4)
            result["fp16" if tag == "f16" else tag] = {
                "grays": round(grays_r, 4), "speedup_vs_f32": round(ms32 / ms_r, 3),
                "rmse_vs_f32_per_mm": rmse,
                "kernels_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof_r.items() if v[1]}}
            if tag == "f16":
                # 16-bit volume and measurements, FP32 corrections
                result["roofline_sweep_kernels"]["f16"] = {
                    "forward": roofline("forward", prof_r["forward"][0] / prof_r["forward"][1],
                                        2 * V + 2 * P + 4 * P, "forward_f16"),
                    "voxel_bp": roofline("voxel_bp", prof_r["voxel_bp"][0] / prof_r["voxel_bp"][1],
                                         4 * P + 2 * 2 * V, "voxel_bp_f16")}
            log(f"{tag.upper()}: {ms_r / 1e3:.3f} s/iter, GRays/s {grays_r:.3f}, speedup {ms32 / ms_r:.3f}, "
                f"rmse vs f32 {rmse:.3e}")
            del meas_r, x_r
        result["fp16"]["target"] = {"speedup_vs_f32": 1.7, "met": result["fp16"]["speedup_vs_f32"] >= 1.7}
        torch.cuda.empty_cache()

    if not args.no_f64:
        # error against the FP64 oracle: FP64, FP32, FP16 and BF16 runs of the
        # same kernels from zero on the same data, per iteration (outside the
        # timed region)
        log(f"accuracy pass: {args.acc_iters} iterations x FP64/FP32/FP16/BF16")
        acc = accuracy(ctx, g, x_true, meas32, args.acc_iters, beta, lam, niter)
        last = acc["iterations"][-1]
        acc["bar"] = {"rmse_vs_f64_per_mm": 1e-3,
                      "met": all(last[k]["rmse_vs_f64"] <= 1e-3 for k in ("f16", "bf16"))}
        result["accuracy_vs_f64"] = acc
        log("vs FP64 after %d: " % args.acc_iters + ", ".join(
            f"{k} rmse {last[k]['rmse_vs_f64']:.3e} rel {last[k]['rel_l2_vs_f64']:.3e}" for k in ("f32", "f16", "bf16")))

    if not args.no_cpu:
        nsv = 32 if args.cpu_full else args.cpu_views
        log(f"CPU baseline: {nsv} sweep views{' (full plan)' if args.cpu_full else ''}")
        meas_host = meas32[: max(nsv, 4)].cpu().numpy()
        result["cpu_baseline"] = cpu_baseline(g, meas_host, beta, lam, niter, nsv, full=args.cpu_full)
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
