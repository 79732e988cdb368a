"""Helpers of the surrogate parity tests: the FP64 oracle frame of a CVR_CONE_NN handle
(its own pyramid as the cone reads store it, its trained weights and input range), the
sparse pyramid check, and image-error summaries (test infrastructure)."""
from __future__ import annotations

import numpy as np

import oracle_lib as O


def bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def stored_levels(levels, precision):
    """The pyramid levels as the renderer's corner blocks hold them: fp32, or rounded to
    nearest even into fp16 (1) or bf16 (2)."""
    if precision == 1:
        return [a.astype(np.float16).astype(np.float64) for a in levels]
    if precision == 2:
        return [bf16_round(a).astype(np.float64) for a in levels]
    return [a.astype(np.float64) for a in levels]


def oracle_cone_params(wl, light, cone):
    return O.cone_params(wl.grid, light, n_steps=cone.n_steps, n_occ=cone.n_occ_cones,
                         occ_half_deg=float(cone.occ_half_deg), ring_deg=float(cone.ring_deg),
                         sdw_half_deg=float(cone.sdw_half_deg), t0=float(cone.t0), r_sigma=float(cone.r_sigma),
                         sigma0=float(cone.sigma0), apply_occ=cone.apply_occlusion, apply_sdw=cone.apply_shadow)


def nn_oracle_frame(wl, r, cam, light, W, H, stride=1, bf16=True, levels=None):
    """FP64 surrogate frame (render_nn) of trained handle `r` for camera `cam`, every
    `stride`-th pixel.  Returns (image, margin, counts)."""
    p = r.params()
    if levels is None:
        levels = r.pyramid()
    pyr = O.OPyr.wrap(wl.grid, stored_levels(levels, p.cone.density_precision))
    lo, hi = r.nn_range()
    m = np.float32(p.ood_margin)
    lo = (lo - m).astype(np.float64)
    hi = (hi + m).astype(np.float64)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    prm = oracle_cone_params(wl, light, p.cone)
    step = float(p.cone.step_size) if p.cone.step_size > 0 else 0.5
    return O.render_nn(wl.volume, tf, pyr, prm, r.nn_weights().astype(np.float64), cam, W, H, lo, hi, bf16=bf16,
                       step=step, et=float(p.cone.et_threshold), scale=wl.scale, stride=stride)


def check_pyramid_sparse(wl, levels, base, rng, n0=40, n1=20):
    """Level 0 at sampled voxels against the 343-tap definition over the volume, each
    level i against the oracle's step from the GPU's level i-1 (relative 1e-4)."""
    bx, by, bz = base
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    _, va = O._vol_args(wl.volume, wl.scale)
    L = O.lib()
    g0 = levels[0]
    for x, y, z in zip(rng.integers(0, bx, n0), rng.integers(0, by, n0), rng.integers(0, bz, n0)):
        ref = -np.log(1 - L.orc_pyr_level0_at(*va, tf.h, bx, by, bz, 1.0, int(x), int(y), int(z)))
        assert abs(g0[z, y, x] - ref) <= 1e-4 * max(1.0, abs(ref)), (0, x, y, z)
    gx, gy, gz = wl.grid
    for lv in range(1, len(levels)):
        prev = np.ascontiguousarray(1.0 - np.exp(-levels[lv - 1].astype(np.float64)))
        pz, py, px = prev.shape
        nz, ny, nx = levels[lv].shape
        for x, y, z in zip(rng.integers(0, nx, n1), rng.integers(0, ny, n1), rng.integers(0, nz, n1)):
            ref = -np.log(1 - L.orc_pyr_next_at(prev.ctypes.data, px, py, pz, nx, ny, nz, float(2 ** lv), gx, gy, gz,
                                                int(x), int(y), int(z)))
            assert abs(levels[lv][z, y, x] - ref) <= 1e-4 * max(1.0, abs(ref)), (lv, x, y, z)


def image_rmse(img, ref):
    """RMSE over every channel of every pixel, and over RGB of the pixels either image
    covers (alpha > 0): the background and the alpha channel, identical by construction
    between two shadings of one march, do not dilute the second figure."""
    d = np.asarray(img, np.float64) - np.asarray(ref, np.float64)
    cov = (np.asarray(img)[..., 3] > 0) | (np.asarray(ref)[..., 3] > 0)
    whole = float(np.sqrt(np.mean(d ** 2)))
    rgb = float(np.sqrt(np.mean(d[cov][:, :3] ** 2))) if cov.any() else 0.0
    return whole, rgb, int(cov.sum())
