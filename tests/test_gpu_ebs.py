"""Parity of the GPU extinction SAT and the HIP extinction-based shading (CVR_EBS)
against the FP64 oracle."""
import numpy as np
import pytest

import oracle_lib as O
from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 2e-3
SAT_RTOL = 1e-6


def sat_rel_err(g, r):
    return float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-30 + 1e-12 * np.abs(r).max())))


# shapes are (nx, ny, nz); rows of w = nx + 2 cells go through the fused x/y pass at
# ceil(w / 64) rounded up to 2, 3, 4, 5, 6, 8, 12 or 16 cells per lane, wider ones through the
# separate x pass; z columns of d = nz + 2 cells through up to 16 register segments of 48
# cells, longer ones one lane per column
@pytest.mark.parametrize("shape,dtype", [((19, 13, 11), N.U8), ((40, 33, 35), N.U16), ((64, 64, 64), N.U16),
                                         # rows past 1024 cells (the separate x pass, beyond one batch of
                                         # row chunks; then y columns of one segment, or one lane each past
                                         # 16 segments) and long y columns in the fused pass
                                         ((1100, 5, 6), N.U8), ((1030, 780, 2), N.U8), ((6, 500, 5), N.U16),
                                         ((7, 900, 4), N.U16),
                                         # z columns of 16 segments, and beyond them
                                         ((5, 4, 700), N.U8), ((5, 7, 800), N.U8),
                                         # 3, 4, 6, 8, 12 and 16 cells per lane in the fused pass
                                         ((150, 9, 3), N.U16), ((200, 5, 4), N.U8), ((350, 4, 3), N.U16),
                                         ((400, 6, 2), N.U8), ((700, 3, 3), N.U16), ((900, 3, 2), N.U8)])
def test_sat_matches_fp64(shape, dtype):
    if dtype == N.U8:
        vol = N.synth_blobs(shape, dtype, n_blobs=4, seed=8)
        rgb, a = W.seeded_tf(8, 4)
    else:
        vol = N.synth_noise(shape, dtype, octaves=4, seed=9, bands=[0.2, 0.5, 0.8])
        wl = W.banded_noise(n=8)
        rgb, a = wl.rgb_iso, wl.alpha_iso
    r = N.Renderer(N.EBS)
    r.set_volume(vol)
    r.set_transfer_function(rgb, a)
    g = r.get_sat(shape)
    ref = O.sat(vol, O.OTF(rgb, a))
    assert sat_rel_err(g, ref) <= SAT_RTOL


def test_sat_full_256_u16():
    """The EBS config's SAT at full size: 256^3 uint16 -> 258^3."""
    wl = W.banded_noise(n=256)
    r = N.Renderer(N.EBS)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    g = r.get_sat((256, 256, 256))
    ref = O.sat(wl.volume, O.OTF(wl.rgb_iso, wl.alpha_iso))
    err = sat_rel_err(g, ref)
    print("sat rel err", err)
    assert err <= SAT_RTOL
    # the last cell is the total extinction
    assert g[-1, -1, -1] == pytest.approx(O.sat_input(wl.volume, O.OTF(wl.rgb_iso, wl.alpha_iso)).sum(), rel=1e-9)


def ebs_pair(wl, w, h, stride=1, cam=None, **params):
    r = N.Renderer(N.EBS)
    r.set_volume(wl.volume, wl.scale)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    p = r.set_params(**params)
    light = wl.light()
    cam = cam or wl.camera()
    img = r.render(cam, w, h, light=light)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    sat = O.sat(wl.volume, tf)
    diag = float(np.linalg.norm(wl.grid))
    prm = O.ebs_params(p, light)
    if prm.sdw_max_dist <= 0:
        prm.sdw_max_dist = np.float32(0.75) * np.float32(diag)
    ref, m = O.render_ebs(wl.volume, tf, sat, prm, cam, w, h, step=0.5, et=0.99, scale=wl.scale, stride=stride)
    return img, ref, m


def assert_parity(img, ref, m, stride=1, max_over_frac=0.005):
    """Strict max-abs on every pixel whose FP64 march made no near-tie decision; the
    near-tie (straddle) pixels may take the other branch, and those that then differ by
    more than the tolerance must stay rare."""
    s = O.compare(img, ref, m, TOL, stride)
    print(s)
    assert s["max_abs"] <= TOL, s
    assert s["n_over_tol"] <= max_over_frac * s["n_pixels"] + 1, s


@pytest.mark.parametrize("shadow_type", [1, 0])
@pytest.mark.parametrize("sat_layout", [1, 0])
def test_ebs_small(shadow_type, sat_layout):
    """Both SAT layouts the march can read (linear (hi, lo) pairs, corner blocks)."""
    wl = W.banded_noise(n=48, width=64, height=64, seed=12)
    img, ref, m = ebs_pair(wl, 64, 64, shadow_type=shadow_type, sat_layout=sat_layout)
    assert_parity(img, ref, m)
    assert ref[..., :3].max() > 0.05


@pytest.mark.parametrize("shadow_type", [1, 0])
def test_ebs_camera_inside_ragged_frame(shadow_type):
    """A camera inside the volume box with a wide field of view, on a frame whose sides
    are not multiples of the tile size."""
    wl = W.banded_noise(n=48, width=61, height=47, seed=14)
    cam = N.Camera((3.0, -2.0, 5.0), (0.0, 0.0, -40.0), (0.0, 1.0, 0.0), 60.0, 61 / 47)
    img, ref, m = ebs_pair(wl, 61, 47, cam=cam, shadow_type=shadow_type)
    assert_parity(img, ref, m)
    assert ref[..., 3].max() > 0.2


def test_ebs_blobs_ao_only_and_shadow_only():
    wl = W.plain_blobs(n=40, width=48, height=48, seed=13)
    for ao, sd in ((1, 0), (0, 1)):
        img, ref, m = ebs_pair(wl, 48, 48, apply_occlusion=ao, apply_shadow=sd, shadow_type=1)
        assert_parity(img, ref, m)


def test_ebs_disabled_lighting_is_an_error():
    wl = W.plain_blobs(n=16, width=8, height=8, seed=1)
    r = N.Renderer(N.EBS)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    r.set_params(apply_occlusion=0, apply_shadow=0)
    with pytest.raises(N.CvrError) as e:
        r.render(wl.camera(), 8, 8)
    assert e.value.code == N.ERR_INVALID


def test_ebs_full_volume_sparse_pixels():
    """The EBS config's 256^3 uint16 banded noise at 256^2; the oracle renders every 8th pixel."""
    wl = W.banded_noise(n=256, width=256, height=256)
    img, ref, m = ebs_pair(wl, 256, 256, stride=8, shadow_type=1)
    assert_parity(img, ref, m, stride=8)


def test_ebs_full_config_1024():
    """The EBS config at its own frame size: 256^3 uint16 banded noise, directional light,
    1024^2; the oracle renders every 32nd pixel (x and y) on its FP64 SAT.  The two SAT
    layouts give the same frame to fp32 rounding."""
    wl = W.banded_noise(n=256, width=1024, height=1024)
    img, ref, m = ebs_pair(wl, 1024, 1024, stride=32, shadow_type=1)
    assert ref[..., 3].max() > 0.2
    assert_parity(img, ref, m, stride=32)
    img0, _, _ = ebs_pair(wl, 1024, 1024, stride=1024, shadow_type=1, sat_layout=0)
    assert np.abs(img0 - img).max() <= 1e-6
