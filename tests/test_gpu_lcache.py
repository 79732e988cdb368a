"""Parity of the object-space light cache (light_cache = 1 on CVR_EBS and CVR_DOS)
against the FP64 oracle: the cache cells, the march over the GPU's own cache (the
march alone), and the whole frame over the oracle's cache."""
import numpy as np
import pytest

import oracle_lib as O
from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 2e-3


def with_cache(tech, wl, res, **params):
    r = N.Renderer(tech)
    r.set_volume(wl.volume, wl.scale)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    p = r.params()
    for k, v in params.items():
        setattr(p, k, v)
    p.light_cache = 1
    p.light_cache_res[:] = list(res)
    r.set_params(p)
    return r, r.params()


def cache_close(g, ref, tol, frac=0.01):
    d = np.abs(g.astype(np.float64) - ref)
    print("cache max", d.max(), "over tol", int((d > tol).sum()), "of", d.size)
    assert (d > tol).sum() <= frac * d.size + 1, d.max()
    assert d.max() <= 50 * tol


def parity(img, ref, m, max_over_frac=0.005):
    s = O.compare(img, ref, m, TOL)
    print(s)
    assert s["max_abs"] <= TOL, s
    assert s["n_over_tol"] <= max_over_frac * s["n_pixels"] + 1, s


@pytest.mark.parametrize("shadow_type,res", [(1, (24, 20, 16)), (0, (32, 32, 32))])
def test_ebs_light_cache(shadow_type, res):
    wl = W.banded_noise(n=48, width=64, height=64, seed=12)
    r, p = with_cache(N.EBS, wl, res, shadow_type=shadow_type)
    light, cam = wl.light(), wl.camera()
    img = r.render(cam, 64, 64, light=light)
    g = r.light_cache()
    assert g.shape == (res[2], res[1], res[0], 2)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    sat = O.sat(wl.volume, tf)
    prm = O.ebs_params(p, light)
    prm.sdw_max_dist = np.float32(0.75) * np.float32(float(np.linalg.norm(wl.grid)))
    ref = O.lc_ebs(wl.volume.shape[::-1], wl.scale, sat, prm, res)
    cache_close(g, ref, 5e-4)
    assert ref[..., 0].min() < 0.9 and ref[..., 1].min() < 0.9
    # the march alone, over the GPU's cache
    own, m = O.render_lc(wl.volume, tf, g, 1, 1, p.ka, p.kd, cam, 64, 64, scale=wl.scale)
    parity(img, own, m)
    # the whole frame over the oracle's cache
    full, m = O.render_lc(wl.volume, tf, ref, 1, 1, p.ka, p.kd, cam, 64, 64, scale=wl.scale)
    parity(img, full, m)
    assert full[..., 3].max() > 0.3


def test_ebs_light_cache_camera_inside_ragged_frame():
    """The march over the GPU's own cache with the camera inside the volume box and a
    frame whose sides are not multiples of the tile size."""
    wl = W.banded_noise(n=48, width=61, height=47, seed=18)
    r, p = with_cache(N.EBS, wl, (24, 20, 16), shadow_type=1)
    light = wl.light()
    cam = N.Camera((3.0, -2.0, 5.0), (0.0, 0.0, -40.0), (0.0, 1.0, 0.0), 60.0, 61 / 47)
    img = r.render(cam, 61, 47, light=light)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    own, m = O.render_lc(wl.volume, tf, r.light_cache(), 1, 1, p.ka, p.kd, cam, 61, 47, scale=wl.scale)
    parity(img, own, m)
    assert own[..., 3].max() > 0.2


@pytest.mark.parametrize("shadow_type,packing", [(2, 1), (0, 2)])
def test_dos_light_cache(shadow_type, packing):
    wl = W.cone_blobs(n=40, width=64, height=48, seed=15)
    r, p = with_cache(N.DOS, wl, (20, 16, 24), apply_shadow=1, shadow_type=shadow_type)
    p.occ.max_packing = packing
    r.set_params(p)
    light, cam = wl.light(), wl.camera()
    img = r.render(cam, 64, 48, light=light)
    g = r.light_cache()
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    pyr = O.OPyr.wrap(wl.grid, [a.astype(np.float64) for a in r.pyramid()])
    ref = O.lc_dos(wl.volume.shape[::-1], wl.scale, pyr, O.dos_params(p, wl.grid, light), cam, 64, 48, (20, 16, 24),
                   sigma0=p.sigma0)
    cache_close(g, ref, 1e-3)
    own, m = O.render_lc(wl.volume, tf, g, 1, 1, p.ka, p.kd, cam, 64, 48, scale=wl.scale)
    parity(img, own, m)
    full, m = O.render_lc(wl.volume, tf, ref, 1, 1, p.ka, p.kd, cam, 64, 48, scale=wl.scale)
    parity(img, full, m)
    assert full[..., 3].max() > 0.2


def test_light_cache_is_rebuilt_per_frame_and_errors():
    wl = W.cone_blobs(n=32, width=32, height=32, seed=16)
    r, p = with_cache(N.DOS, wl, (8, 8, 8))
    with pytest.raises(N.CvrError):
        r.light_cache()  # nothing rendered yet
    r.render(W.orbit_camera(wl.grid, 32, 32, 10.0), 32, 32, light=wl.light())
    a = r.light_cache()
    r.render(W.orbit_camera(wl.grid, 32, 32, 190.0), 32, 32, light=wl.light())
    b = r.light_cache()
    assert np.abs(a[..., 0] - b[..., 0]).max() > 1e-3  # the occlusion cone follows the eye
    p = r.params()
    p.light_cache_res[:] = [0, 8, 8]
    with pytest.raises(N.CvrError):
        r.set_params(p)
    e = N.Renderer(N.EBS)
    q = e.params()
    q.light_cache = 1
    q.light_cache_res[:] = [8, 600, 8]
    with pytest.raises(N.CvrError):
        e.set_params(q)
