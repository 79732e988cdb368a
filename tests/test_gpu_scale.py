"""Parity of the GPU multi-scale output resampling (cvr_resample_frame) against the FP64
oracle, for every mode and kernel, on ragged sizes (the digital filter's short-sequence
path included), and on a rendered frame."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W

pytestmark = pytest.mark.gpu


def run(frame, dw, dh, mode, kernel, in_place=False):
    src = torch.from_numpy(np.ascontiguousarray(frame, dtype=np.float32)).cuda()
    dst = torch.zeros((dh, dw, 4), dtype=torch.float32, device="cuda")
    work = None
    if mode == N.SCALE_UPSCALE and kernel >= N.FILTER_CARDINAL_BSPLINE3:
        work = src if in_place else torch.empty_like(src)
    N.resample_frame(src, dst, mode, kernel, work=work)
    torch.cuda.synchronize()
    return dst.cpu().numpy(), src.cpu().numpy()


CASES = [(N.SCALE_MULTISAMPLE, 64, 48, 32, 24), (N.SCALE_DOWNSCALE, 66, 50, 33, 25), (N.SCALE_DOWNSCALE, 96, 72, 24, 18),
         (N.SCALE_UPSCALE, 33, 25, 66, 50), (N.SCALE_UPSCALE, 20, 15, 80, 60), (N.SCALE_UPSCALE, 7, 5, 21, 15)]


@pytest.mark.parametrize("kernel", range(6))
@pytest.mark.parametrize("mode,sw,sh,dw,dh", CASES)
def test_resample_matches_fp64(mode, sw, sh, dw, dh, kernel):
    f = np.random.default_rng(sw * 7 + kernel).uniform(0, 1, (sh, sw, 4))
    got, src_after = run(f, dw, dh, mode, kernel)
    ref = O.resample(f.astype(np.float32), dw, dh, mode, kernel)
    err = np.abs(got - ref).max()
    scale = max(1.0, np.abs(ref).max())
    print(mode, kernel, err)
    assert err <= 2e-5 * scale
    assert np.array_equal(src_after, f.astype(np.float32))  # the source is left alone


def test_upscale_in_place_filters_the_source():
    f = np.random.default_rng(9).uniform(0, 1, (24, 32, 4))
    got, src_after = run(f, 64, 48, N.SCALE_UPSCALE, N.FILTER_CARDINAL_BSPLINE3, in_place=True)
    ref = O.resample(f.astype(np.float32), 64, 48, N.SCALE_UPSCALE, N.FILTER_CARDINAL_BSPLINE3)
    assert np.abs(got - ref).max() <= 2e-5
    assert not np.array_equal(src_after, f.astype(np.float32))


def test_render_at_twice_the_screen_then_downscale():
    wl = W.plain_blobs(n=64, width=128, height=96, seed=51)
    fw, fh = N.multiscale_size(128, 96, 2, 2)
    assert (fw, fh) == (256, 192)
    assert N.multiscale_size(128, 96, -2, -3) == (64, 32)
    r = N.Renderer(N.RC1P)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    cam = W.orbit_camera(wl.grid, fw, fh, 30.0, 20.0)
    frame = torch.zeros((fh, fw, 4), dtype=torch.float32, device="cuda")
    r.render(cam, fw, fh, out=frame)
    screen = torch.zeros((96, 128, 4), dtype=torch.float32, device="cuda")
    N.resample_frame(frame, screen, N.SCALE_DOWNSCALE, N.FILTER_MITCHELL_NETRAVALI)
    torch.cuda.synchronize()
    ref = O.resample(frame.cpu().numpy(), 128, 96, N.SCALE_DOWNSCALE, N.FILTER_MITCHELL_NETRAVALI)
    assert np.abs(screen.cpu().numpy() - ref).max() <= 2e-5
    assert ref[..., 3].max() > 0.3


def test_resample_errors():
    src = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    dst = torch.zeros((16, 16, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(N.CvrError):
        N.resample_frame(src, dst, N.SCALE_UPSCALE, N.FILTER_CARDINAL_OMOMS3)  # needs a work buffer
    with pytest.raises(N.CvrError):
        N.resample_frame(src, dst, 7, N.FILTER_HAT)
    with pytest.raises(N.CvrError):
        N.multiscale_size(128, 96, 0, 2)
