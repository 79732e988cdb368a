"""cvr_render_noalloc: the allocation-free surrogate render (tile marks in the library's
fixed per-GPU array) gives the pixels cvr_render gives, call after call, for frames of
decreasing size and with tiles that take the second (exact-tracer) pass; it rejects what
it cannot do without allocating."""
import ctypes as C

import numpy as np
import pytest
import torch

from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prepared():
    wl = W.cone_blobs(n=48, width=96, height=80, seed=23)
    r = N.Renderer(N.CONE_NN)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    p = r.params()
    p.n_train, p.n_val, p.epochs, p.batch = 1 << 16, 1 << 12, 2, 4096
    p.cone.n_levels = 6
    p.rmse_gate = 1.0  # the network renders (two passes over the tile marks), not the exact tracer
    r.set_params(p)
    r.prepare()
    assert r.nn_report().exact_fallback == 0
    return wl, r


def frames(r, cam, light, w, h):
    ref = torch.zeros((h, w, 4), device="cuda")
    r.render(cam, w, h, light=light, out=ref)
    got = torch.full((h, w, 4), -1.0, device="cuda")
    r.render_noalloc(cam, w, h, got, light=light)
    torch.cuda.synchronize()
    return got.cpu().numpy(), ref.cpu().numpy()


def test_noalloc_equals_render_on_two_calls_the_second_smaller(prepared):
    wl, r = prepared
    light = wl.light()
    for (w, h, az) in ((96, 80, 30.0), (40, 24, 75.0)):
        cam = W.orbit_camera(wl.grid, w, h, az, 20.0)
        got, ref = frames(r, cam, light, w, h)
        print(w, h, float(np.abs(got - ref).max()))
        assert ref[..., 3].max() > 0.2
        assert np.array_equal(got, ref)


def test_noalloc_with_tiles_marked_for_the_second_pass(prepared):
    """A narrowed input range sends some samples to the exact tracer: the first pass marks
    their tiles in the fixed array, the second renders them again and clears the marks;
    repeated calls (and a smaller frame after a larger one) match cvr_render."""
    wl, r = prepared
    p = r.params()
    margin = p.ood_margin
    try:
        r.set_params(ood_margin=-0.02)
        light = wl.light()
        for (w, h) in ((96, 80), (96, 80), (56, 40)):
            cam = W.orbit_camera(wl.grid, w, h, 30.0, 20.0)
            got, ref = frames(r, cam, light, w, h)
            assert np.array_equal(got, ref), (w, h)
        r.collect_stats(True)
        r.render(cam, 56, 40, light=light)
        st = r.stats()
        assert 0 < st.fallback_samples < st.shaded_samples
    finally:
        r.collect_stats(False)
        r.set_params(ood_margin=margin)


def test_noalloc_rejects_what_needs_allocation(prepared):
    wl, r = prepared
    cam, light = wl.camera(), wl.light()
    L = N.lib()
    # a frame with more workgroup tiles than the fixed array holds (4 x 16-pixel wave tiles:
    # 513 x 128 workgroup tiles), checked before any work is queued
    big = torch.zeros((1, 1, 4), device="cuda")
    rc = L.cvr_render_noalloc(r.h, C.byref(cam), C.byref(light), 4104, 4096, C.c_void_p(big.data_ptr()), None)
    assert rc == N.ERR_INVALID and b"limit" in L.cvr_last_error()
    # host memory
    host = np.zeros((80, 96, 4), np.float32)
    rc = L.cvr_render_noalloc(r.h, C.byref(cam), C.byref(light), 96, 80, C.c_void_p(host.ctypes.data), None)
    assert rc == N.ERR_INVALID
    # an unprepared handle (pyramid, cone tables and network not built yet)
    fresh = N.Renderer(N.CONE_NN)
    fresh.set_volume(wl.volume)
    fresh.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    out = torch.zeros((80, 96, 4), device="cuda")
    with pytest.raises(N.CvrError) as e:
        fresh.render_noalloc(cam, 96, 80, out, light=light)
    assert e.value.code == N.ERR_STATE
    # another technique
    plain = N.Renderer(N.RC1P)
    plain.set_volume(wl.volume)
    plain.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    with pytest.raises(N.CvrError) as e:
        plain.render_noalloc(cam, 96, 80, out)
    assert e.value.code == N.ERR_UNSUPPORTED
    # the handle still renders afterwards
    got, ref = frames(r, cam, light, 96, 80)
    assert np.array_equal(got, ref)
