"""Parity of the HIP directional-occlusion renderer with cone section tables (CVR_DOS)
and of the ray-sampled cone ground truth (CVR_CONE_GT) against the FP64 oracle."""
import numpy as np
import pytest

import oracle_lib as O
from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 2e-3


def make(tech, wl, **params):
    r = N.Renderer(tech)
    r.set_volume(wl.volume, wl.scale)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    p = r.params()
    for k, v in params.items():
        if k in ("occ", "sdw"):
            for kk, vv in v.items():
                setattr(getattr(p, k), kk, vv)
        elif k == "base_res":
            p.base_res[:] = list(v)
        else:
            setattr(p, k, v)
    r.set_params(p)
    return r, r.params()


def assert_parity(img, ref, m, stride=1, max_over_frac=0.002):
    s = O.compare(img, ref, m, TOL, stride)
    print(s)
    assert s["max_abs"] <= TOL, s
    assert s["n_over_tol"] <= max_over_frac * s["n_pixels"] + 1, s
    return s


# ---------------------------------------------------------------------------- DOS
@pytest.mark.parametrize("occ,sdw", [({}, {}), ({"half_deg": 30.0, "max_packing": 2}, {"half_deg": 5.0, "max_packing": 1}),
                                     ({"half_deg": 200.0, "max_packing": 2, "d_sigma": 9.0, "r_sigma": 7.0,
                                       "covered_dist": 3.5}, {})])
def test_dos_sections_match_oracle(occ, sdw):
    wl = W.cone_blobs(n=48, width=8, height=8, seed=21)
    r, p = make(N.DOS, wl, occ=occ, sdw=sdw)
    want = O.dos_params(p, wl.grid)
    for which, spec in ((0, want.occ), (1, want.sdw)):
        rows, counts, axes, adj7 = r.dos_sections(which)
        ref, rc, ra, radj = O.dos_sections(spec, p.sigma0)
        assert list(counts) == list(rc)
        assert rows.shape[0] == ref.shape[0]
        assert np.allclose(rows, ref[:, :4], rtol=2e-7, atol=1e-7)
        assert np.allclose(axes, ra, atol=1e-7)
        assert abs(adj7 - radj) <= 1e-7


@pytest.mark.parametrize("shadow_type,apply_occ,packing", [(0, 1, 1), (2, 1, 2), (1, 0, 1), (0, 0, 0)])
def test_dos_render_small_own_pyramid(shadow_type, apply_occ, packing):
    wl = W.cone_blobs(n=40, width=64, height=48, seed=15)
    r, p = make(N.DOS, wl, apply_occlusion=apply_occ, apply_shadow=1, shadow_type=shadow_type, base_res=(32, 40, 24),
                spot_max_deg=30.0, occ={"max_packing": packing}, sdw={"half_deg": 3.0, "max_packing": packing})
    light = wl.light()
    img = r.render(wl.camera(), 64, 48, light=light)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    pyr = O.OPyr.build(wl.volume, tf, base=(32, 40, 24))
    ref, m = O.render_dos(wl.volume, tf, pyr, O.dos_params(p, wl.grid, light), wl.camera(), 64, 48)
    assert_parity(img, ref, m)
    assert ref[..., 3].max() > 0.2


def test_dos_render_camera_inside_ragged_frame():
    """A camera inside the volume box with a wide field of view (occlusion cones towards
    an eye among the samples), on a frame whose sides are not multiples of the tile size."""
    wl = W.cone_blobs(n=40, width=61, height=47, seed=17)
    r, p = make(N.DOS, wl, apply_occlusion=1, apply_shadow=1, shadow_type=0, base_res=(32, 40, 24))
    light = wl.light()
    cam = N.Camera((3.0, -2.0, 5.0), (0.0, 0.0, -40.0), (0.0, 1.0, 0.0), 60.0, 61 / 47)
    img = r.render(cam, 61, 47, light=light)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    pyr = O.OPyr.build(wl.volume, tf, base=(32, 40, 24))
    ref, m = O.render_dos(wl.volume, tf, pyr, O.dos_params(p, wl.grid, light), cam, 61, 47)
    assert_parity(img, ref, m)
    assert ref[..., 3].max() > 0.2


def test_dos_render_anisotropic_u16():
    vol = N.synth_noise((36, 28, 20), N.U16, octaves=3, seed=7, bands=[0.3, 0.6])
    rgb, a = W.seeded_tf(17, 5, max_density=4095)  # a 12-bit table over the 16-bit range
    r = N.Renderer(N.DOS)
    scale = (1.0, 1.5, 2.0)
    r.set_volume(vol, scale)
    r.set_transfer_function(rgb, a, max_density=4095)
    p = r.params()
    p.apply_shadow, p.shadow_type, p.step_size = 1, 2, 0.625
    p.base_res[:] = [40, 32, 24]
    r.set_params(p)
    grid = (36 * 1.0, 28 * 1.5, 20 * 2.0)
    cam = W.orbit_camera(grid, 40, 40, 50.0, 30.0)
    light = W.light_from_direction((0.3, 0.8, -0.4), grid)
    img = r.render(cam, 40, 40, light=light)
    tf = O.OTF(rgb, a, max_density=4095)
    pyr = O.OPyr.build(vol, tf, base=(40, 32, 24), scale=scale)
    ref, m = O.render_dos(vol, tf, pyr, O.dos_params(p, grid, light), cam, 40, 40, step=0.625, scale=scale)
    assert_parity(img, ref, m)
    assert ref[..., 3].max() > 0.1


def test_dos_full_defaults_sparse():
    """256^3 u8 with 64 blobs at 1024^2 with the reference defaults (128^3 pyramid,
    occlusion cone of 20 degrees packed up to 3) plus the point-light shadow cone; the
    oracle renders every 32nd pixel on the GPU's pyramid."""
    wl = W.cone_blobs(n=256)
    r, p = make(N.DOS, wl, apply_shadow=1)
    light = wl.light()
    img = r.render(wl.camera(), 1024, 1024, light=light)
    pyr = O.OPyr.wrap(wl.grid, [a.astype(np.float64) for a in r.pyramid()])
    assert pyr.levels == 8
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    ref, m = O.render_dos(wl.volume, tf, pyr, O.dos_params(p, wl.grid, light), wl.camera(), 1024, 1024, stride=32)
    assert_parity(img, ref, m, stride=32)


def test_dos_parameter_errors():
    wl = W.cone_blobs(n=16, width=8, height=8)
    r = N.Renderer(N.DOS)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    p = r.params()
    p.apply_occlusion = p.apply_shadow = 0
    with pytest.raises(N.CvrError):
        r.set_params(p)
    p = r.params()
    p.shadow_type = 3
    with pytest.raises(N.CvrError):
        r.set_params(p)
    p = r.params()
    p.occ.initial_step = 0.0
    with pytest.raises(N.CvrError):
        r.set_params(p)
    p = r.params()  # sigma would have to grow without bound: the schedule never closes
    p.occ.half_deg, p.occ.max_packing, p.occ.d_sigma, p.occ.r_sigma = 60.0, 0, 3.0, 0.5
    with pytest.raises(N.CvrError):
        r.set_params(p)
    with pytest.raises(N.CvrError):
        r.dos_sections(2)


# ---------------------------------------------------------------------------- cone ground truth
def test_gt_rays_match_oracle():
    wl = W.cone_blobs(n=16, width=8, height=8)
    r, p = make(N.CONE_GT, wl, occ_rays=12, occ_aperture_deg=60.0, sdw_rays=5, sdw_aperture_deg=4.0)
    occ, sdw = O.gt_rays(O.gt_params(p, wl.grid))
    assert np.allclose(r.gt_rays(0), occ, atol=1e-7)
    assert np.allclose(r.gt_rays(1), sdw, atol=1e-7)


@pytest.mark.parametrize("shadow_type,n_occ,n_sdw", [(0, 8, 2), (2, 3, 1), (1, 4, 1)])
def test_gt_render_small(shadow_type, n_occ, n_sdw):
    wl = W.cone_blobs(n=40, width=48, height=40, seed=16)
    r, p = make(N.CONE_GT, wl, occ_rays=n_occ, occ_aperture_deg=60.0, sdw_rays=n_sdw, sdw_aperture_deg=5.0,
                shadow_type=shadow_type)
    light = wl.light()
    img = r.render(wl.camera(), 48, 40, light=light)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    ref, m = O.render_gt(wl.volume, tf, O.gt_params(p, wl.grid, light), wl.camera(), 48, 40)
    assert_parity(img, ref, m)
    assert ref[..., 3].max() > 0.2


def test_gt_render_128_sparse():
    wl = W.cone_blobs(n=128, width=256, height=256, seed=3)
    r, p = make(N.CONE_GT, wl, occ_rays=8, occ_aperture_deg=30.0, sdw_rays=1, sdw_aperture_deg=10.0)
    light = wl.light()
    img = r.render(wl.camera(), 256, 256, light=light)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    ref, m = O.render_gt(wl.volume, tf, O.gt_params(p, wl.grid, light), wl.camera(), 256, 256, stride=8)
    assert_parity(img, ref, m, stride=8)
