"""Precision report: each output of the HIP path against the FP64 oracle, per storage
precision of the sampled densities (run on an MI355X: python tests/precision_report.py).

Images: max-abs per RGBA channel over non-straddle pixels, RMSE, pixels over 2e-3, on
every `stride`-th pixel of the full-size configurations.  SAT: relative error of the
stored (hi, lo) table.  Pyramid: relative error of the fp32 levels.  MLP: bf16-MFMA
inference against the FP64 network on the same weights.  Surrogate: validation RMSE of
the trained network against the exact tracer.  Prints a markdown table.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle_lib as O  # noqa: E402
import surrogate_oracle as SO  # noqa: E402
from cpp_volume_rendering_amd import native as N  # noqa: E402
from cpp_volume_rendering_amd import workloads as W  # noqa: E402

ROWS = []


def row(output, storage, s):
    ROWS.append((output, storage, s))
    print(f"| {output} | {storage} | {s} |", flush=True)


def img_stats(img, ref, m, stride):
    c = O.compare(img, ref, m, 2e-3, stride)
    rgba = ", ".join(f"{x:.1e}" for x in c["max_abs_rgba"])
    return (f"max-abs (r,g,b,a) {rgba}; RMSE {c['rmse']:.1e}; over 2e-3: {c['n_over_tol']}/{c['n_pixels']} "
            f"(straddles {c['n_straddle']})")


def main():
    print("# Precision report: the HIP path against the FP64 oracle\n")
    print("Command (from the repository root, on an MI355X): `python tests/precision_report.py > profiles/precision_report.md`\n")
    print("| output | storage / arithmetic | error vs FP64 oracle |")
    print("|---|---|---|")
    # plain caster, 128^3 u8 at 512^2 (every 4th pixel)
    wl = W.plain_blobs()
    r = N.Renderer(N.RC1P)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    img = r.render(wl.camera(), 512, 512)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    ref, m = O.render_rc1p(wl.volume, tf, wl.camera(), 512, 512, stride=4)
    row("rc1pass image 128³ u8 @512²", "u8 voxels, fp32 march", img_stats(img, ref, m, 4))

    # SAT and EBS, 256^3 u16 at 1024^2
    wl = W.banded_noise()
    r = N.Renderer(N.EBS)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    r.set_params(shadow_type=1)
    sat = r.get_sat((256, 256, 256))
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    sref = O.sat(wl.volume, tf)
    rel = np.abs(sat - sref) / np.maximum(np.abs(sref), 1e-30)
    row("SAT 258³", "x, y sums in fp64, z in fp32 two-float; stored (hi, lo)", f"max relative {rel[sref > 0].max():.1e}")
    light = wl.light()
    img = r.render(wl.camera(), 1024, 1024, light=light)
    prm = O.ebs_params(r.params(), light)
    if prm.sdw_max_dist <= 0:
        prm.sdw_max_dist = np.float32(0.75) * np.float32(float(np.linalg.norm(wl.grid)))
    ref, m = O.render_ebs(wl.volume, tf, sref, prm, wl.camera(), 1024, 1024, stride=32)
    row("EBS image 256³ u16 @1024²", "fp32 march, FP64 box sums", img_stats(img, ref, m, 32))

    # cone tracer and DOS, 256^3 u8 at 1024^2, per pyramid storage
    wl = W.cone_blobs()
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    light = wl.light()
    names = {0: "fp32", 1: "fp16", 2: "bf16"}
    for prec in (0, 1, 2):
        r = N.Renderer(N.CONE_TRACE)
        r.set_volume(wl.volume)
        r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
        r.set_params(density_precision=prec)
        img = r.render(wl.camera(), 1024, 1024, light=light)
        lv = r.pyramid()
        if prec == 0:
            small = W.cone_blobs(n=48, seed=15)
            rs = N.Renderer(N.CONE_TRACE)
            rs.set_volume(small.volume)
            rs.set_transfer_function(small.rgb_iso, small.alpha_iso)
            g = rs.pyramid()
            pr = O.OPyr.build(small.volume, O.OTF(small.rgb_iso, small.alpha_iso), nlev=6)
            worst = max(float(np.max(np.abs(g[l] - pr.level(l)) / np.maximum(1.0, np.abs(pr.level(l)))))
                        for l in range(pr.levels))
            row("extinction pyramid 48³, 6 levels", "fp32", f"max relative (floor 1) {worst:.1e}")
        pyr = O.OPyr.wrap(wl.grid, [a.astype(np.float64) for a in lv])
        ref, m = O.render_cone(wl.volume, tf, pyr, O.cone_params(wl.grid, light), wl.camera(), 1024, 1024, stride=32)
        row("cone tracer image 256³ @1024²", f"pyramid {names[prec]}, fp32 cones", img_stats(img, ref, m, 32))
        d = N.Renderer(N.DOS)
        d.set_volume(wl.volume)
        d.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
        p = d.set_params(apply_shadow=1, density_precision=prec)
        img = d.render(wl.camera(), 1024, 1024, light=light)
        dpyr = O.OPyr.wrap(wl.grid, [a.astype(np.float64) for a in d.pyramid()])
        ref, m = O.render_dos(wl.volume, tf, dpyr, O.dos_params(p, wl.grid, light), wl.camera(), 1024, 1024, stride=32)
        row("DOS image 256³ @1024² (128³ pyramid)", f"pyramid {names[prec]}, fp32 sections", img_stats(img, ref, m, 32))

    # surrogate at the headline size, 512^3 at 1024^2: the frame against the FP64 surrogate
    # frame (render_nn over the same trained weights, bf16 rounding points included) per
    # pyramid storage, the image against the exact tracer, the bf16 MLP against the FP64
    # network, and the trained network against the exact tracer
    wl = W.surrogate_blobs_noise()
    cam, light = wl.camera(), wl.light()
    for prec in (1, 0, 2):
        r = N.Renderer(N.CONE_NN)
        r.set_volume(wl.volume)
        r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
        p = r.params()
        p.cone.density_precision = prec
        r.set_params(p)
        r.prepare()
        img = r.render(cam, 1024, 1024, light=light)
        ref, m, _ = SO.nn_oracle_frame(wl, r, cam, light, 1024, 1024, stride=32)
        row("surrogate image 512³ @1024²", f"pyramid {names[prec]}, bf16 MLP{' (default)' if prec == 1 else ''}",
            img_stats(img, ref, m, 32) + " (vs the FP64 surrogate frame)")
        if prec != 1:
            continue
        ct = N.Renderer(N.CONE_TRACE)
        ct.set_volume(wl.volume)
        ct.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
        ct.set_params(r.params().cone)
        whole, rgb, ncov = SO.image_rmse(img, ct.render(cam, 1024, 1024, light=light))
        row("surrogate image 512³ @1024² vs the exact cone tracer", "pyramid fp16, bf16 MLP",
            f"image RMSE {whole:.4f} over the frame, {rgb:.4f} over RGB of {ncov} covered pixels (acceptance: 0.01)")
        w = r.nn_weights().astype(np.float64)
        rng = np.random.default_rng(3)
        X = rng.uniform(0, 1, (4096, 16)).astype(np.float32)
        y = r.nn_eval(X, path=1)
        yr = O.mlp_forward(w, X.astype(np.float64))[:, :2]
        row("surrogate MLP outputs (16-64-64-64-16)", "bf16 operands, fp32 accumulate",
            f"max-abs {np.abs(y - yr).max():.1e} (FP64 network, same weights)")
        rep = r.nn_report()
        row("surrogate occlusion / shadow", "bf16 MLP vs exact cone tracer",
            f"validation RMSE {rep.val_rmse_occ:.4f} / {rep.val_rmse_sdw:.4f} "
            f"(network-free estimate {rep.base_rmse_occ:.4f} / {rep.base_rmse_sdw:.4f})")


if __name__ == "__main__":
    main()
