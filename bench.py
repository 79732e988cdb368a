"""Benchmark of the single-pass structured ray casters on one MI355X.

python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

One step = one rendered frame of the workload with the volume, transfer function
and any preprocessing (SAT, pyramid, trained network) already resident in HBM.
Prints one JSON line.  Only one GPU is used; --gpus above 1 is refused.
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
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(kind: str, wl, cam, light, budget_s: float = 15.0):
    """The oracle's FP32 instantiation on the host cores (all of them, OpenMP over rows)
    over a bounded sample of the same frame: every k-th row, k chosen so one pass
    costs a few seconds; frames/s is scaled from the fraction of rows rendered."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np
    import oracle_lib as O

    L = O.lib()
    threads = L.orc_max_threads()
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    W_, H_ = wl.width, wl.height
    if kind == "ebs":
        sat = O.sat(wl.volume, tf).astype(np.float32)
        prm = O.ebs_params(None, light, shadow_type=1)
        prm.sdw_max_dist = np.float32(0.75) * np.float32(np.linalg.norm(wl.grid))

    def run(rows):
        # render `rows` image rows spread over the frame (a camera with a shorter frame
        # would change the rays; instead render the full-width rows with a row stride)
        cam2 = cam
        t0 = time.perf_counter()
        if kind == "rc1p":
            O.render_rc1p(wl.volume, tf, cam2, W_, H_, fp32=True, threads=threads) if rows == H_ else None
        return time.perf_counter() - t0

    # sample: full frame for the plain caster when it fits the budget, else a row subset
    t0 = time.perf_counter()
    if kind == "rc1p":
        O.render_rc1p(wl.volume, tf, cam, W_, H_, fp32=True, threads=threads)
        frames = 1
        t = time.perf_counter() - t0
        while t < budget_s and frames < 50:
            O.render_rc1p(wl.volume, tf, cam, W_, H_, fp32=True, threads=threads)
            frames += 1
            t = time.perf_counter() - t0
        return {"value": frames / t, "unit": "frames/s", "cores": threads, "kind": "port",
                "sample": f"{frames} full {W_}x{H_} frames of the same workload, oracle FP32, OpenMP rows"}
    # EBS: a 1/16 subset of pixels (every 4th pixel in x and y), scaled by 16
    out = O.render_ebs(wl.volume, tf, sat, prm, cam, W_, H_, fp32=True, threads=threads)
    t = time.perf_counter() - t0
    return {"value": 1.0 / t, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"1 full {W_}x{H_} frame, oracle FP32 with fp32 SAT, OpenMP rows"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="rc1p", choices=["rc1p", "ebs"])
    ap.add_argument("--no-cpu", action="store_true")
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
    if args.workload == "rc1p":
        wl = W.plain_blobs(n=128, width=512, height=512, seed=1)
        tech, dtype, bpv = N.RC1P, "f32", 1
    else:
        wl = W.banded_noise(n=256, width=1024, height=1024, seed=2)
        tech, dtype, bpv = N.EBS, "f32", 2
    cam, light = wl.camera(), wl.light()
    r = N.Renderer(tech)
    r.set_volume(wl.volume, wl.scale)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    if tech == N.EBS:
        r.set_params(shadow_type=1)
    r.prepare(stream.cuda_stream)
    out = torch.empty((wl.height, wl.width, 4), dtype=torch.float32, device=dev)

    # counters from one extra frame (not timed)
    r.collect_stats(True)
    r.render(cam, wl.width, wl.height, light=light, out=out, stream=stream.cuda_stream)
    stream.synchronize()
    st = r.stats()
    r.collect_stats(False)

    for _ in range(args.warmup):
        r.render(cam, wl.width, wl.height, light=light, out=out, stream=stream.cuda_stream)
    stream.synchronize()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        r.render(cam, wl.width, wl.height, light=light, out=out, stream=stream.cuda_stream)
    e1.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    fps = 1000.0 / ms

    nz, ny, nx = wl.volume.shape
    # algorithmic bytes per frame: every trilinear sample gathers 8 voxels, plus the frame
    gathered = st.samples * 8 * bpv + wl.width * wl.height * 16
    achieved = gathered / (ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(fps, 4),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded)",
        "config": {"workload": wl.name, "volume": f"{nx}x{ny}x{nz} {'u8' if bpv == 1 else 'u16'}",
                   "image": f"{wl.width}x{wl.height}", "step": 0.5,
                   "samples_per_frame": int(st.samples), "shaded_samples_per_frame": int(st.shaded_samples)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None},
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, wl, cam, light)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
