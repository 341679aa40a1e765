#!/usr/bin/env python3
"""Benchmark: fwd+back GRays/s and seconds per SART-FISTA-TV iteration at
512^3 voxels / 1024 views of 1024^2, FP32 storage vs FP16 storage.

Workload (synthetic, seeded): 512^3 volume of 0.5 mm voxels holding the
10-ellipsoid phantom scaled to +-100 mm; 1024 views over 360 degrees onto a
1024x1024 detector of 0.4 mm pixels, SID 1000 mm / SDD 1500 mm; Poisson noise
with I0 = 1e5 photons per ray, then the log transform.  One step = one outer
iteration: a SART sweep over all 1024 views in the reference's view-by-view
order (relaxation 2/32, the reference's 2/Lip_con), TV-FGP with 20 inner
iterations (lambda 1e-3) and the FISTA momentum, all device resident.
Inputs are in HBM before timing starts.

value = fwd+back GRays/s of the FP32 path = rays per sweep / (time of the
Siddon correction kernel + time of the voxel-driven update kernel), both timed
live with HIP events on the context stream.  The same run reports the FP16
path, seconds per iteration of both, the roofline of the dominant kernel,
FDK (Hann) over the same projections, the autograd pair (ray sums + matched
adjoint) on a sample of views, and the CPU oracle (OpenMP restatement of the
reference kernels) on a bounded sample of the same sweep.

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
    the FP64 path (same data, same schedule) and against the phantom."""
    import torch
    paths = {"f64": torch.float64, "f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}
    state = {}
    for k, dt in paths.items():
        meas = meas32 if dt in (torch.float32, torch.float64) else meas32.to(dt)
        state[k] = [torch.zeros(g.vol_shape, device="cuda", dtype=dt), torch.zeros(g.vol_shape, device="cuda", dtype=dt),
                    1.0, meas, 0.0]
    out = {"workload": f"512^3 / 1024 views SART(subset 1, beta {beta}) + TV-FGP({niter}, {lam}) + FISTA",
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
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(g, meas_host, nviews, beta):
    """Oracle SART correction + voxel update for the first `nviews` views of the
    same sweep, FP32, OpenMP over all host cores."""
    import numpy as np
    from oracle import oracle as orc
    threads = orc.num_threads()
    v = np.zeros(g.vol_shape, dtype=np.float32)
    t0 = time.perf_counter()
    for j in range(nviews):
        c = orc.sart_correction(g, v, meas_host[j:j + 1], view0=j, nviews=1, prec=32)
        v = orc.backproject_voxel(g, v, c, view0=j, nviews=1, beta=beta, mode=1, prec=32)
    dt = time.perf_counter() - t0
    rays = nviews * g.nrows * g.ncols
    return {"value": rays / dt / 1e9, "unit": "GRays/s", "cores": threads, "kind": "port",
            "sample": f"first {nviews} views of the 512^3 / 1024-view SART sweep (correction + voxel update), "
                      f"FP32, {dt:.1f} s, OpenMP {threads} threads"}


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
    ap.add_argument("--cpu-views", type=int, default=8)
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 reference iterations (accuracy)")
    ap.add_argument("--accuracy", type=int, default=0,
                    help="accuracy mode: N SART-FISTA-TV iterations in FP64 (the GPU FP64 instantiation as the "
                         "oracle), FP32, FP16 and BF16; prints per-iteration RMSE / relative L2 against FP64 and "
                         "the phantom as one JSON line, then exits")
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
    # relaxation 2/Lip_con = 2/32, the reference's SART-FISTA step
    # (V63/main_recon.cu:58,161); FGP argument lambda = 1e-3
    beta, lam, niter = 0.0625, 1e-3, 20
    ctx = Context(0)
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
        # correction walks ray by ray (the views above take the column walk)
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

    if args.accuracy:
        return accuracy(ctx, g, x_true, meas32, args.accuracy, beta, lam, niter)

    log(f"FP32 path: warmup {args.warmup}, steps {args.steps}")
    ms32, prof32, x32 = run_path(ctx, g, meas32, torch.float32, args.steps, args.warmup, beta, lam, niter)
    fwd_ms, fwd_n = prof32["forward"]
    bp_ms, bp_n = prof32["voxel_bp"]
    rays_per_sweep = g.rays
    grays = rays_per_sweep * args.steps / ((fwd_ms + bp_ms) * 1e-3) / 1e9
    log(f"FP32: {ms32 / 1e3:.3f} s/iter, forward {fwd_ms / fwd_n:.3f} ms x {fwd_n}, "
        f"voxel bp {bp_ms / bp_n:.3f} ms x {bp_n}, GRays/s {grays:.3f}")

    V, P = g.nvox, g.nrows * g.ncols
    kernels = {
        "forward": (fwd_ms / fwd_n, 4 * V + 4 * P + 4 * P),     # read f, read b, write c
        "voxel_bp": (bp_ms / bp_n, 4 * P + 2 * 4 * V),           # read c, read+write f
    }
    dom = max(kernels, key=lambda k: prof32[k][0])
    avg_ms, abytes = kernels[dom]
    achieved = abytes / (avg_ms * 1e-3) / 1e9
    # HBM bytes and VALU instructions per launch of the same kernels, from the
    # committed rocprofv3 counter passes (they cannot be read from inside this
    # process); the live time turns the VALU count into an issue-rate fraction
    traffic, valu_frac = None, None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        tj = json.loads(tf.read_text())
        traffic = tj.get("hbm_bytes_per_launch", {}).get(dom)
        vi = tj.get("valu_insts_per_launch", {}).get(dom)
        if vi:
            valu_frac = round(vi / (VALU_ISSUE_PER_S * avg_ms * 1e-3), 4)

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
            "sart": "view by view (subset 1), relaxation 0.0625 (2/Lip_con)",
            "tv_fgp": {"inner_iterations": niter, "lambda": lam},
            "rays_per_sweep": rays_per_sweep,
        },
        "s_per_iter": {"f32": round(ms32 / 1e3, 4)},
        "kernels_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof32.items() if v[1]},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": abytes, "valu_issue_frac": valu_frac},
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
    result["beam_styles"] = style_figures(ctx, g, x32)
    log("beam styles: " + ", ".join(f"{k} {v['grays']} GRays/s" for k, v in result["beam_styles"].items()
                                     if isinstance(v, dict)))
    result["multisource_pair"] = multisource_figures(ctx)
    log(f"multi-source pair: {result['multisource_pair']['grays']} GRays/s")
    result["io"] = io_figures(ctx, g, meas32, x32)
    log("files: " + ", ".join(f"{k} {v}" for k, v in result["io"].items() if k != "note"))

    if not args.no_fp16:
        log("FP16 path")
        meas16 = meas32.to(torch.float16)
        ms16, prof16, x16 = run_path(ctx, g, meas16, torch.float16, args.steps, args.warmup, beta, lam, niter)
        f16 = prof16["forward"][0] + prof16["voxel_bp"][0]
        grays16 = rays_per_sweep * args.steps / (f16 * 1e-3) / 1e9
        diff = (x16.float() - x32).double()
        rmse = float(torch.sqrt((diff * diff).mean()))
        result["s_per_iter"]["f16"] = round(ms16 / 1e3, 4)
        result["fp16"] = {"grays": round(grays16, 4), "speedup_vs_f32": round(ms32 / ms16, 3),
                          "rmse_vs_f32_per_mm": rmse,
                          "kernels_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof16.items() if v[1]}}
        log(f"FP16: {ms16 / 1e3:.3f} s/iter, GRays/s {grays16:.3f}, speedup {ms32 / ms16:.3f}, rmse {rmse:.3e}")
        del meas16

    if not args.no_f64:
        # error against the FP64 oracle: the FP64 instantiation of the same
        # kernels, run for the same number of iterations on the same data
        # (outside the timed region)
        log(f"FP64 reference: {args.warmup + args.steps} iterations")
        x64 = torch.zeros(g.vol_shape, device="cuda", dtype=torch.float64)
        xp64 = torch.zeros_like(x64)
        t64 = 1.0
        for _ in range(args.warmup + args.steps):
            t64 = ctx.sart_fista_iteration(g, x64, xp64, meas32, t64, beta=beta, subset=1, fgp_lambda=lam,
                                           fgp_niter=niter)
        torch.cuda.synchronize()
        ref_n = float(torch.linalg.vector_norm(xp64))
        acc = {"iterations": args.warmup + args.steps, "reference": "FP64 GPU instantiation of the same kernels"}
        for k, xk in (("f32", x32), ("f16", x16 if not args.no_fp16 else None)):
            if xk is None:
                continue
            d = xk.double() - xp64
            acc[k] = {"rmse_per_mm": float(torch.sqrt((d * d).mean())),
                      "rel_l2": float(torch.linalg.vector_norm(d)) / ref_n}
        result["accuracy_vs_f64"] = acc
        log(f"vs FP64: " + ", ".join(f"{k} rmse {v['rmse_per_mm']:.3e} rel {v['rel_l2']:.3e}"
                                     for k, v in acc.items() if isinstance(v, dict)))
        del x64, xp64

    if not args.no_cpu:
        log(f"CPU baseline: {args.cpu_views} views")
        meas_host = meas32[: args.cpu_views].cpu().numpy()
        result["cpu_baseline"] = cpu_baseline(g, meas_host, args.cpu_views, beta)
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
