"""Benchmark of the single-pass structured ray casters on one MI355X.

python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

Default workload (the headline metric): 512^3 uint8 (64 seeded Gaussian blobs plus
4-octave value noise), 1024^2 image, directional occlusion + shadow.  One step = one
frame of the learned-surrogate renderer (CVR_CONE_NN) with volume, pyramid and
trained network resident in HBM; the exact cone-traced renderer (CVR_CONE_TRACE) is
timed the same way beside it and judges the surrogate's image RMSE.  Other
workloads: rc1p (128^3 plain caster at 512^2), ebs (256^3 extinction-based shading
at 1024^2), cone (256^3 exact cone tracer at 1024^2), dos (256^3 section-table
directional occlusion + point-light shadow at 1024^2), batch (4 uneven volumes x 16
views x 4 lights, 512^2, one graph-captured persistent launch, work shape and step
schedule autotuned), sat (the EBS config's summed-area-table build alone), ebs-lc and
dos-lc (the 256^3 configs through a 32^3 object-space light cache).  Prints one JSON
line; the headline exits with status 3 (after printing) when the surrogate's image
RMSE against the exact tracer exceeds 0.01.  Each gather-bound line is priced at the
gather peak of the cache level its read footprint fits, measured by a microbenchmark
first.  Only one GPU is used; --gpus above 1 is refused.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "frames/s 512³@1024² DAO+shadow: neural vs cone-traced at RMSE≤0.01; HBM GB/s"
WORKLOAD_METRIC = {
    "rc1p": "frames/s 128³ u8@512² plain single-pass compositing (rc1pass); HBM GB/s",
    "ebs": "frames/s 256³ u16@1024² extinction-based shading (SAT AO + directional cone shadow); HBM GB/s",
    "cone": "frames/s 256³ u8@1024² cone-traced DAO+shadow (8+1 cones x 64 steps, 6-level pyramid); HBM GB/s",
    "dos": "frames/s 256³ u8@1024² section-table directional occlusion + point-light shadow; HBM GB/s",
    "ebs-lc": "frames/s 256³ u16@1024² extinction-based shading through a 32³ object-space light cache; HBM GB/s",
    "dos-lc": "frames/s 256³ u8@1024² directional occlusion + shadow through a 32³ object-space light cache; HBM GB/s",
}
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
MLP_FLOP_PER_SAMPLE = 2 * (16 * 64 + 64 * 64 + 64 * 64 + 64 * 16)  # 20480
# image RMSE (RGB over covered pixels) the batch's step search may add against the
# default step: half of the 0.01 the surrogate is held to against the exact tracer
STEP_RMSE_GATE = 0.005


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_frames(r, wl, cam, light, out, stream, steps, warmup):
    import torch
    for _ in range(warmup):
        r.render(cam, wl.width, wl.height, light=light, out=out, stream=stream.cuda_stream)
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        r.render(cam, wl.width, wl.height, light=light, out=out, stream=stream.cuda_stream)
    e1.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def stats_frame(r, wl, cam, light, out, stream):
    r.collect_stats(True)
    r.render(cam, wl.width, wl.height, light=light, out=out, stream=stream.cuda_stream)
    stream.synchronize()
    st = r.stats()
    r.collect_stats(False)
    return st


def gather_bytes(tech, st, r, bpv):
    """Algorithmic bytes one frame gathers: each trilinear read counts its 8 texels (volume
    samples at the voxel size, pyramid reads as fp32 or 16-bit values); the cone tracer
    reads 2 levels per step, the section-table renderer one level per section and
    sub-cone, the surrogate 16 whole-level reads per shaded sample."""
    from cpp_volume_rendering_amd import native as N
    p = r.params()
    vol = st.samples * 8 * bpv
    if tech == N.EBS:
        # each box sum reads 8 trilinear SAT taps of 8 float2 (hi, lo) cells
        return vol + st.box_sums * 8 * 8 * 8
    if tech == N.DOS:
        b = 4 if p.density_precision == 0 else 2
        reads = 0
        for which, on in ((0, p.apply_occlusion), (1, p.apply_shadow)):
            if on:
                c = r.dos_sections(which)[1]
                reads += int(c[0]) + 3 * int(c[1]) + 7 * int(c[2])
        return vol + st.shaded_samples * reads * 8 * b
    if tech == N.CONE_TRACE:
        b = 4 if p.density_precision == 0 else 2
        return vol + st.shaded_samples * (p.n_occ_cones + 1) * p.n_steps * 2 * 8 * b
    if tech == N.CONE_NN:
        b = 4 if p.cone.density_precision == 0 else 2
        exact = st.fallback_samples * (p.cone.n_occ_cones + 1) * p.cone.n_steps * 2 * 8 * b
        return vol + st.shaded_samples * 16 * 8 * b + exact
    return vol


# Working sets of the gather-peak probe: each fits one level of the memory hierarchy
# (L1 per CU, L2 per XCD (4 MB), the shared 256 MB MALL) or none (HBM only).
PROBE_SETS = (("l1", 16 << 10), ("l2", 2 << 20), ("mall", 64 << 20), ("hbm", 4 << 30))
LEVEL_CAPACITY = (("l2", 4 << 20), ("mall", 256 << 20), ("hbm", 1 << 62))
FASTER = {"hbm": "mall", "mall": "l2", "l2": "l1"}


def gather_peaks(N, stream):
    """Measured ceilings (GB/s of independent 32-byte record reads) of each memory level."""
    return {k: round(N.probe_gather(v, reps=3, stream=stream.cuda_stream), 1) for k, v in PROBE_SETS}


def footprint_level(nbytes):
    for name, cap in LEVEL_CAPACITY:
        if nbytes <= cap:
            return name
    return "hbm"


def gather_roofline(achieved, footprint, peaks, traffic, ms):
    """Roofline of a gather-bound kernel.  Its reads come from the level that holds their
    footprint (the bytes of every structure the kernel reads); HBM is priced at its
    published 8 TB/s, a cache level at its measured gather ceiling.  A kernel whose rate
    exceeds that level's ceiling is served by a faster level (locality), so the first
    faster level whose ceiling is above the achieved rate bounds it instead."""
    level = footprint_level(footprint)
    peak = HBM_PEAK_GBS if level == "hbm" else peaks[level]
    bound = level
    while achieved > peak and bound in FASTER:
        bound = FASTER[bound]
        peak = peaks[bound]
    roof = {"bound": bound, "achieved": round(achieved, 2), "peak": peak, "unit": "GB/s",
            "frac": round(achieved / peak, 5), "traffic": traffic,
            "footprint_bytes": int(footprint), "footprint_level": level, "gather_peaks_gbs": peaks}
    if traffic:
        # secondary figure: the measured L2-miss traffic (FETCH_SIZE + WRITE_SIZE) per frame
        # against the published HBM peak
        roof["traffic_gbs"] = round(traffic / (ms * 1e-3) / 1e9, 1)
        roof["traffic_frac_of_hbm_peak"] = round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return roof


def read_footprint(tech, r, wl, bpv, light_cache):
    """Bytes of the structures the dominant kernel reads: the corner-block volume (8
    values per voxel), the SAT's 64-byte corner blocks, the pyramid's corner blocks
    (8 values per cell at the stored width), the frame written."""
    from cpp_volume_rendering_amd import native as N
    nz, ny, nx = wl.volume.shape
    total = nx * ny * nz * 8 * bpv + wl.width * wl.height * 16
    if light_cache:
        return total + 32 ** 3 * 8
    if tech == N.EBS:
        # the (hi, lo) SAT as the march reads it: linear 8-byte cells, or 64-byte corner blocks
        total += (nx + 2) * (ny + 2) * (nz + 2) * (8 if r.params().sat_layout == 1 else 64)
    if tech in (N.DOS, N.CONE_TRACE, N.CONE_NN):
        p = r.params()
        # DOS and the cone tracer carry density_precision themselves, the surrogate in its cone params
        prec = p.cone.density_precision if tech == N.CONE_NN else p.density_precision
        cells = sum(int(a.size) for a in r.pyramid())
        total += cells * 8 * (4 if prec == 0 else 2)
    return total


def measured_traffic(kernel, workload, width):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size), when one was
    measured on this workload."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_{width}.md"
    if not f.exists():
        return None
    import re
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", f.read_text())
    return int(m.group(1)) if m else None


def host_cpu():
    """nproc and the lscpu model name of the host running the CPU baseline."""
    import subprocess
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.lower().startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count() or 0, model


def median_of_5(run, arg):
    """One warm-up call, then the median wall time of 5 calls."""
    run(arg)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        run(arg)
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[2], ts


def cpu_baseline(kind, wl, cam, light, params=None, budget_s=18.0):
    """The oracle's FP32 instantiation on all host cores (OpenMP over rows) on a bounded
    sample of the same frame: every `stride`-th pixel in x and y (whole frames for the
    plain caster), one warm-up then the median of 5 runs, frames/s scaled by the fraction
    of pixels rendered (stride^2)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np
    import oracle_lib as O

    threads = O.lib().orc_max_threads()
    nproc, model = host_cpu()
    host = {"cores": threads, "nproc": nproc, "cpu_model": model, "kind": "port", "protocol": "1 warm-up, median of 5"}
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    W_, H_ = wl.width, wl.height
    if kind == "rc1p":
        def frames_run(k):
            for _ in range(k):
                O.render_rc1p(wl.volume, tf, cam, W_, H_, fp32=True, threads=threads)
        t0 = time.perf_counter()
        frames_run(1)
        per = max(1e-3, time.perf_counter() - t0)
        k = max(1, int(budget_s / 6 / per))
        med, ts = median_of_5(frames_run, k)
        return {"value": k / med, "unit": "frames/s", **host,
                "runs_s": [round(t, 3) for t in ts],
                "sample": f"runs of {k} full {W_}x{H_} frames, oracle FP32, OpenMP over rows"}
    if kind in ("ebs-lc", "dos-lc"):
        # the cache (FP64 oracle) is part of every frame; the march over it in FP32
        if kind == "ebs-lc":
            sat64 = O.sat(wl.volume, tf)  # preprocessing, not part of a frame
            prm = O.ebs_params(None, light, shadow_type=1)
            prm.sdw_max_dist = np.float32(0.75) * np.float32(np.linalg.norm(wl.grid))
            t0 = time.perf_counter()
            cache = O.lc_ebs(wl.volume.shape[::-1], wl.scale, sat64, prm, (32, 32, 32), threads=threads)
            t_cache = time.perf_counter() - t0
        else:
            t0 = time.perf_counter()
            p = params["native"]
            cache = O.lc_dos(wl.volume.shape[::-1], wl.scale, params["pyramid"], O.dos_params(p, wl.grid, light), cam,
                             W_, H_, (32, 32, 32), sigma0=p.sigma0, threads=threads)
            t_cache = time.perf_counter() - t0
        what = f"FP64 oracle light cache ({t_cache:.2f} s per frame) + FP32 march over it"

        def run(stride):
            O.render_lc(wl.volume, tf, cache, 1, 1, 0.5, 0.5, cam, W_, H_, fp32=True, threads=threads, stride=stride)
    elif kind == "ebs":
        sat = O.sat(wl.volume, tf).astype(np.float32)
        prm = O.ebs_params(None, light, shadow_type=1)
        prm.sdw_max_dist = np.float32(0.75) * np.float32(np.linalg.norm(wl.grid))
        what = "oracle FP32 with fp32 SAT"

        def run(stride):
            O.render_ebs(wl.volume, tf, sat, prm, cam, W_, H_, fp32=True, threads=threads, stride=stride)
    elif kind == "dos":
        pyr = params["pyramid"]
        prm = O.dos_params(params["native"], wl.grid, light)
        what = "section-table cone integration, oracle FP32"

        def run(stride):
            O.render_dos(wl.volume, tf, pyr, prm, cam, W_, H_, fp32=True, threads=threads, stride=stride)
    else:
        # cone tracer: the exact tracer's FP32 port on the GPU-built pyramid (pyramid build not timed)
        pyr = params["pyramid"]
        prm = O.cone_params(wl.grid, light)
        what = "exact cone tracer, oracle FP32"

        def run(stride):
            O.render_cone(wl.volume, tf, pyr, prm, cam, W_, H_, fp32=True, threads=threads, stride=stride)
    # a coarse probe, then a pixel stride sized so that one run takes about a sixth of
    # `budget_s` (the work scales with the number of pixels, 1 / stride^2); then one
    # warm-up and 5 timed runs at that stride
    target = budget_s / 6.0
    stride = 64
    t0 = time.perf_counter()
    run(stride)
    t = time.perf_counter() - t0
    for _ in range(3):
        if t >= 0.5 * target or stride <= 1:
            break
        # round up: near 1 a truncated stride would overshoot the budget (stride 1.6 -> 1 is
        # 2.6x the pixels)
        stride = max(1, min(stride - 1, math.ceil(stride * (t / target) ** 0.5)))
        t0 = time.perf_counter()
        run(stride)
        t = time.perf_counter() - t0
    med, ts = median_of_5(run, stride)
    frame_s = med * stride * stride
    if kind in ("ebs-lc", "dos-lc"):
        frame_s += t_cache
    return {"value": 1.0 / frame_s, "unit": "frames/s", **host, "runs_s": [round(x, 3) for x in ts],
            "sample": f"one pixel in {stride} along x and along y of one {W_}x{H_} frame (1/{stride * stride} of it, "
                      f"median {med:.2f} s), {what}, OpenMP over rows; scaled to whole frames"}


def bench_batch(args, N, W, torch, stream):
    """Batched render: 4 uneven volumes x 16 views x 4 lights = 256 images of 512^2 with
    the surrogate, one graph-captured persistent launch per step, shape autotuned first."""
    t0 = time.perf_counter()
    wls, jobs = W.batch_scene(scale_down=max(1, args.size))
    rs = []
    for wl in wls:
        r = N.Renderer(N.CONE_NN)
        r.set_volume(wl.volume, wl.scale)
        r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
        r.prepare(stream.cuda_stream)
        rs.append(r)
    stream.synchronize()
    log(f"batch: {len(wls)} volumes prepared (pyramids, training) in {time.perf_counter() - t0:.1f}s")
    b = N.Batch(512, 512)
    for r in rs:
        b.add_volume(r)
    b.set_jobs(jobs)
    out = torch.empty((len(jobs), 512, 512, 4), dtype=torch.float32, device="cuda")
    t1 = time.perf_counter()
    tune = b.autotune(out, stride=8, stream=stream.cuda_stream, max_step_rmse=STEP_RMSE_GATE)
    log(f"autotune {time.perf_counter() - t1:.1f}s: step_scale {tune.step_scale} (image RMSE {tune.step_rmse:.5f} "
        f"against the default step, gate {STEP_RMSE_GATE}) tile_w {tune.tile_w} shade_batch {tune.shade_batch} "
        f"queues {tune.queues} wg/cu {tune.workgroups_per_cu} span {tune.span} ({tune.ms:.2f} ms on every 8th job)")
    for _ in range(args.warmup):
        b.render(out, stream.cuda_stream)
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        b.render(out, stream.cuda_stream)
    e1.record(stream)
    stream.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    st = b.stats()
    ips = len(jobs) / (ms * 1e-3)
    flops = st.shaded_samples * MLP_FLOP_PER_SAMPLE
    achieved = flops / (ms * 1e-3) / 1e12
    line = {
        "metric": "images/s, batched surrogate DAO+shadow (4 volumes x 16 views x 4 lights, 512^2)",
        "value": round(ips, 3), "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic (seeded)",
        "config": {"workload": "batch_4vol_16view_4light_512", "images_per_step": len(jobs),
                   "volumes": [f"{wl.name} {wl.volume.shape[::-1]}" for wl in wls],
                   "tuning": {"tile_w": tune.tile_w, "shade_batch": round(tune.shade_batch, 3),
                              "queues": tune.queues, "workgroups_per_cu": tune.workgroups_per_cu, "span": tune.span,
                              "step_scale": round(tune.step_scale, 3),
                              "step_rmse_vs_default_step": round(tune.step_rmse, 6),
                              "step_rmse_gate": STEP_RMSE_GATE},
                   "samples_per_step": int(st.samples), "shaded_samples_per_step": int(st.shaded_samples),
                   "fallback_samples_per_step": int(st.fallback_samples)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / BF16_PEAK_TFLOPS, 6), "traffic": None},
    }
    print(json.dumps(line), flush=True)


def bench_sat(args, N, W, stream):
    """The extinction SAT build of the EBS config (256^3 u16 -> 258^3 two-float cells) on
    its own: one step = one build (a transfer-function upload between steps invalidates
    the table), device time of the prefix passes from events around them.
    Algorithmic bytes: the voxels read once, the (hi, lo) table written once."""
    wl = W.banded_noise(n=args.size or 256, width=1024, height=1024, seed=2)
    r = N.Renderer(N.EBS)
    r.set_volume(wl.volume, wl.scale)
    r.collect_stats(True)
    scan, blocks = [], []
    for i in range(args.warmup + args.steps):
        r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
        r.prepare(stream.cuda_stream)
        st = r.stats()
        if i >= args.warmup:
            scan.append(st.sat_scan_ms)
            blocks.append(st.sat_blocks_ms)
    ms = sum(scan) / len(scan)
    ms_b = sum(blocks) / len(blocks)
    nz, ny, nx = wl.volume.shape
    algo = nx * ny * nz * 2 + (nx + 2) * (ny + 2) * (nz + 2) * 8
    achieved = algo / (ms * 1e-3) / 1e9
    log(f"SAT build {nx}x{ny}x{nz} u16: prefix passes {ms:.4f} ms, corner blocks {ms_b:.4f} ms")
    line = {
        "metric": "SAT builds/s 256³ u16 -> 258³ two-float extinction prefix sums; HBM GB/s",
        "value": round(1000.0 / ms, 3), "unit": "builds/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (two-float)", "data": "synthetic (seeded)",
        "config": {"workload": wl.name, "volume": f"{nx}x{ny}x{nz} u16", "sat": f"{nx + 2}x{ny + 2}x{nz + 2} float2",
                   "algorithmic_bytes": algo, "corner_block_copy_ms": round(ms_b, 5),
                   "kernels": "k_sat_xy (x and y), k_sat_seg (z)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="nn",
                    choices=["nn", "cone", "rc1p", "ebs", "dos", "ebs-lc", "dos-lc", "batch", "sat"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--size", type=int, default=0, help="override the volume size (tests)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with three 256 MiB device-to-device copies (calibrates rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from cpp_volume_rendering_amd import native as N
    from cpp_volume_rendering_amd import workloads as W

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    if args.workload in ("batch", "sat"):
        if args.workload == "batch":
            bench_batch(args, N, W, torch, stream)
        else:
            bench_sat(args, N, W, stream)
        if args.calib_copy:
            calib_copy(torch)
        return
    peaks = gather_peaks(N, stream)
    log(f"gather peaks (GB/s of 32-byte reads): {peaks}")
    t_setup = time.perf_counter()
    extra = {}
    if args.workload == "rc1p":
        wl = W.plain_blobs(n=args.size or 128, width=512, height=512, seed=1)
        tech, dtype, bpv = N.RC1P, "f32", 1
    elif args.workload in ("ebs", "ebs-lc"):
        wl = W.banded_noise(n=args.size or 256, width=1024, height=1024, seed=2)
        tech, dtype, bpv = N.EBS, "f32", 2
    elif args.workload in ("dos", "dos-lc"):
        wl = W.cone_blobs(n=args.size or 256, width=1024, height=1024, seed=3)
        tech, dtype, bpv = N.DOS, "f32", 1
    elif args.workload == "cone":
        wl = W.cone_blobs(n=args.size or 256, width=1024, height=1024, seed=3)
        tech, dtype, bpv = N.CONE_TRACE, "f32", 1
    else:
        wl = W.surrogate_blobs_noise(n=args.size or 512, width=1024, height=1024, seed=4)
        tech, dtype, bpv = N.CONE_NN, "bf16", 1
    log(f"workload {wl.name}: volume {wl.volume.shape} ready in {time.perf_counter() - t_setup:.1f}s")
    cam, light = wl.camera(), wl.light()
    r = N.Renderer(tech)
    r.set_volume(wl.volume, wl.scale)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    light_cache = args.workload.endswith("-lc")
    if tech == N.EBS:
        r.set_params(shadow_type=1, light_cache=int(light_cache))
    if tech == N.DOS:
        r.set_params(apply_shadow=1, light_cache=int(light_cache))
    t_prep = time.perf_counter()
    r.prepare(stream.cuda_stream)
    stream.synchronize()
    prep_s = time.perf_counter() - t_prep
    log(f"prepare (pyramid / SAT / training) {prep_s:.1f}s")
    out = torch.empty((wl.height, wl.width, 4), dtype=torch.float32, device=dev)

    st = stats_frame(r, wl, cam, light, out, stream)
    ms = time_frames(r, wl, cam, light, out, stream, args.steps, args.warmup)
    fps = 1000.0 / ms
    log(f"{wl.name}: {ms:.3f} ms/frame, {fps:.2f} frames/s, samples {st.samples}, shaded {st.shaded_samples}, "
        f"fallback {st.fallback_samples}")

    nz, ny, nx = wl.volume.shape
    config = {"workload": wl.name, "volume": f"{nx}x{ny}x{nz} {'u8' if bpv == 1 else 'u16'}",
              "image": f"{wl.width}x{wl.height}", "step": 0.5,
              "samples_per_frame": int(st.samples), "shaded_samples_per_frame": int(st.shaded_samples)}
    if tech == N.CONE_NN:
        rep = r.nn_report()
        p = r.params()
        ct = N.Renderer(N.CONE_TRACE)
        ct.set_volume(wl.volume, wl.scale)
        ct.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
        ct.set_params(p.cone)
        ct.prepare(stream.cuda_stream)
        out_ct = torch.empty_like(out)
        st_ct = stats_frame(ct, wl, cam, light, out_ct, stream)
        ms_ct = time_frames(ct, wl, cam, light, out_ct, stream, max(2, args.steps // 4), 1)
        r.render(cam, wl.width, wl.height, light=light, out=out, stream=stream.cuda_stream)
        stream.synchronize()
        rmse = float(torch.sqrt(torch.mean((out - out_ct) ** 2)).item())
        # RGB over the pixels either frame covers: alpha and the background are the same
        # in both by construction and would dilute the figure
        cov = (out[..., 3] > 0) | (out_ct[..., 3] > 0)
        d = (out - out_ct)[cov][:, :3]
        rmse_rgb = float(torch.sqrt(torch.mean(d.double() ** 2)).item()) if d.numel() else 0.0
        gate_ok = max(rmse, rmse_rgb) <= 0.01
        log(f"exact cone tracer: {ms_ct:.3f} ms/frame; surrogate speedup {ms_ct / ms:.2f}x, image RMSE {rmse:.5f} "
            f"(RGB over {int(cov.sum().item())} covered pixels {rmse_rgb:.5f}): {'pass' if gate_ok else 'FAIL'}")
        config.update({"pyramid_levels": p.cone.n_levels, "cones": "8 occlusion @30 deg + 1 shadow @10 deg, 64 steps",
                       "mlp": "16-64-64-64-16 ReLU, bf16 MFMA, fp32 accumulate",
                       "train_samples": p.n_train, "train_steps": rep.steps, "prepare_s": round(prep_s, 2),
                       "fallback_samples_per_frame": int(st.fallback_samples),
                       "val_rmse_occ": round(rep.val_rmse_occ, 5), "val_rmse_sdw": round(rep.val_rmse_sdw, 5),
                       "baseline_rmse_occ": round(rep.base_rmse_occ, 5), "baseline_rmse_sdw": round(rep.base_rmse_sdw, 5),
                       "exact_cone_fps": round(1000.0 / ms_ct, 4), "exact_cone_ms": round(ms_ct, 4),
                       "speedup_vs_exact": round(ms_ct / ms, 3), "image_rmse_vs_exact": round(rmse, 6),
                       "image_rmse_rgb_covered_vs_exact": round(rmse_rgb, 6),
                       "rmse_gate": "pass (<= 0.01)" if gate_ok else "FAIL (> 0.01)",
                       "exact_gate_fallback": int(rep.exact_fallback)})
        # the network's share of the frame: 20480 FLOP per shaded sample
        config["mlp_tflops"] = round(st.shaded_samples * MLP_FLOP_PER_SAMPLE / (ms * 1e-3) / 1e12, 3)
    # roofline of the dominant kernel (one launch per frame): the algorithmic bytes of its
    # gathers, every trilinear read counted as its 8 texels, plus the frame written once
    if light_cache:
        # the march over the cache: volume taps + 8 float2 cache taps per shaded sample
        gathered = st.samples * 8 * bpv + st.shaded_samples * 8 * 8 + wl.width * wl.height * 16
    else:
        gathered = gather_bytes(tech, st, r, bpv) + wl.width * wl.height * 16
    achieved = gathered / (ms * 1e-3) / 1e9
    config["gathered_bytes_per_frame"] = int(gathered)
    kernel = {N.RC1P: "k_rc1p", N.EBS: "k_ebs", N.DOS: "k_dos_render", N.CONE_TRACE: "k_cone_render",
              N.CONE_NN: "k_nn_render"}[tech]
    if light_cache:
        kernel = "k_lc_render"
        config["light_cache"] = "32x32x32 (occlusion, shadow) rebuilt every frame"
    config["kernel"] = kernel
    roof = gather_roofline(achieved, read_footprint(tech, r, wl, bpv, light_cache), peaks,
                           measured_traffic(kernel, wl.name, wl.width), ms)
    line = {
        "metric": WORKLOAD_METRIC.get(args.workload, METRIC), "value": round(fps, 4), "unit": "frames/s",
        "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded)", "config": config, "roofline": roof,
    }
    if not args.no_cpu:
        kind = {N.RC1P: "rc1p", N.EBS: "ebs", N.DOS: "dos"}.get(tech, "cone")
        if light_cache:
            kind += "-lc"
        params = None
        if kind in ("cone", "dos", "dos-lc"):
            sys.path.insert(0, str(ROOT / "tests"))
            import oracle_lib as O
            levels = (r if tech in (N.CONE_TRACE, N.DOS) else ct).pyramid()
            params = {"pyramid": O.OPyr.wrap(wl.grid, [a.astype(np.float64) for a in levels]), "native": r.params()}
        t0 = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(kind, wl, cam, light, params)
        log(f"cpu baseline {time.perf_counter() - t0:.1f}s: {line['cpu_baseline']}")
    print(json.dumps(line), flush=True)
    if args.calib_copy:
        calib_copy(torch)
    if tech == N.CONE_NN and not gate_ok:
        log("the surrogate's image RMSE against the exact tracer exceeds 0.01: the headline line is not valid")
        sys.exit(3)


def calib_copy(torch):
    """Three copies of 256 MiB: a kernel of known traffic (read 256 MiB, write 256 MiB)
    to scale the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    n = 256 << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"calibration copies: 3 x {n} bytes read + written")


if __name__ == "__main__":
    main()
