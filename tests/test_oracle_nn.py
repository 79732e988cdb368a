"""Pins the oracle's surrogate pieces on the CPU: the FP64 network forward against a
numpy forward, its bf16-rounding-point variant against a numpy restatement with bit-level
rounding, and the surrogate-shaded frame (render_nn) against the plain ray caster on
pyramids whose features have closed forms, and against the exact cone renderer when
every sample falls outside the network's range."""
import math

import numpy as np
import pytest

import oracle_lib as O
from cpp_volume_rendering_amd.workloads import plain_blobs

LAYERS = [(64, 16), (64, 64), (64, 64), (16, 64)]


def bf16_np(a):
    """fp32 round-to-nearest-even to bfloat16 by bit arithmetic, as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def np_forward(prm, X, rnd=None):
    a = X if rnd is None else rnd(X)
    o = 0
    for l, (out, inp) in enumerate(LAYERS):
        Wm = prm[o:o + out * inp].reshape(out, inp)
        o += out * inp
        b = prm[o:o + out]
        o += out
        if rnd is not None:
            Wm = rnd(Wm)
        z = a @ Wm.T + b
        if l < 3:
            a = np.maximum(z, 0)
            if rnd is not None:
                a = rnd(a)
        else:
            a = z
    return a


def test_mlp_forward_matches_numpy():
    rng = np.random.default_rng(11)
    P = O.lib().orc_mlp_nparams()
    prm = rng.normal(0, 0.3, P)
    X = rng.uniform(-0.5, 1.5, (200, 16))
    Y = O.mlp_forward(prm, X)
    ref = np_forward(prm, X)
    assert np.max(np.abs(Y - ref)) <= 1e-12 * (1 + np.abs(ref).max())
    assert (ref[:, :2] != 0).all()


def test_mlp_forward_bf16_matches_numpy_rounding():
    rng = np.random.default_rng(12)
    P = O.lib().orc_mlp_nparams()
    prm = rng.normal(0, 0.3, P)
    X = rng.uniform(0, 1, (200, 16))
    Y = O.mlp_forward_bf16(prm, X)
    ref = np_forward(prm, X, bf16_np)
    assert np.max(np.abs(Y - ref)) <= 1e-12 * (1 + np.abs(ref).max())
    # the rounding points matter: the bf16 network is not the FP64 one
    assert np.max(np.abs(Y - O.mlp_forward(prm, X))) > 1e-4


def small_scene(n=24):
    wl = plain_blobs(n=n, width=n, height=n, seed=4)
    return wl, O.OTF(wl.rgb_iso, wl.alpha_iso), wl.camera()


def output_bias_net(b0, b1):
    """Zero weights, zero hidden biases, output biases (b0, b1): the network outputs
    (b0, b1) whatever its inputs."""
    P = O.lib().orc_mlp_nparams()
    w = np.zeros(P)
    w[P - 16], w[P - 15] = b0, b1
    return w


@pytest.mark.parametrize("bf16", [True, False])
@pytest.mark.parametrize("b,apply,m", [
    ((-0.3, -0.6), (1, 1), 0.5 * 0.7 + 0.5 * 0.4),   # Io = 0.7, Is = 0.4
    ((0.5, -2.0), (1, 1), 0.5),                      # clamped: Io = 1.5 -> 1, Is = -1 -> 0
    ((-0.9, -0.25), (0, 1), 0.75),                   # occlusion off: ka = 0, m = Is
    ((-0.2, -0.7), (1, 0), 0.8),                     # shadow off: m = Io
])
def test_render_nn_on_a_zero_pyramid_scales_the_plain_frame(bf16, b, apply, m):
    """A pyramid of zeros makes every input 0, the network-free estimate (1, 1) and the
    network's output its output biases, so (Io, Is) = clamp(1 + b); each shaded sample's
    colour is then scaled by (Io ka + Is kd) / (ka + kd) and the frame is the plain ray
    caster's with its RGB times that factor and the same alpha."""
    wl, tf, cam = small_scene()
    pyr = O.OPyr.wrap(wl.grid, [np.zeros((24 >> i,) * 3) for i in range(4)])
    prm = O.cone_params(wl.grid, wl.light(), apply_occ=apply[0], apply_sdw=apply[1])
    img, _, cnt = O.render_nn(wl.volume, tf, pyr, prm, output_bias_net(*b), cam, 24, 24, bf16=bf16)
    ref, _ = O.render_rc1p(wl.volume, tf, cam, 24, 24)
    assert np.allclose(img[..., 3], ref[..., 3], atol=1e-13)
    assert np.allclose(img[..., :3], m * ref[..., :3], atol=1e-13)
    assert cnt[0] > 0 and cnt[1] == 0
    assert ref[..., 3].max() > 0.1


def test_render_nn_uniform_medium_baseline_closed_form():
    """A uniform extinction tau everywhere and cones short enough to stay inside the box:
    the inputs are the segment opacities 1 - exp(-tau (b - a)), so the network-free
    estimate is exp(-tau (occ_max - t0)) for occlusion (axial and ring alike) and
    exp(-tau (sdw_max - t0)) for the shadow; a zero network leaves it unchanged."""
    n, tau = 64, 0.02
    vol = np.zeros((n, n, n), np.uint8)
    vol[28:36, 26:38, 27:37] = 200
    tf = O.OTF([[1.0, 0.6, 0.3, 0], [0.2, 0.9, 0.5, 255]], [[0.0, 0], [0.5, 255]])
    grid = (float(n),) * 3
    pyr = O.OPyr.wrap(grid, [np.full((n >> i,) * 3, tau) for i in range(6)])
    light = (0.3, 0.8, -0.5)
    prm = O.cone_params(grid, occ_max=10.0, sdw_max=14.0, light_fwd=light)
    from cpp_volume_rendering_amd.workloads import orbit_camera
    cam = orbit_camera(grid, 32, 32)
    img, _, cnt = O.render_nn(vol, tf, pyr, prm, np.zeros(O.lib().orc_mlp_nparams()), cam, 32, 32)
    ref, _ = O.render_rc1p(vol, tf, cam, 32, 32)
    m = 0.5 * math.exp(-tau * 9.0) + 0.5 * math.exp(-tau * 13.0)
    assert cnt[0] > 0
    assert np.allclose(img[..., 3], ref[..., 3], atol=1e-13)
    assert np.allclose(img[..., :3], m * ref[..., :3], atol=1e-12)


def test_render_nn_outside_the_range_is_the_exact_cone_renderer():
    """An empty input range sends every shaded sample to the exact tracer: the frame is the
    cone renderer's; the full range sends none."""
    wl, tf, cam = small_scene(28)
    pyr = O.OPyr.build(wl.volume, tf, nlev=4)
    prm = O.cone_params(wl.grid, wl.light())
    w = np.random.default_rng(5).normal(0, 0.2, O.lib().orc_mlp_nparams())
    img, _, cnt = O.render_nn(wl.volume, tf, pyr, prm, w, cam, 28, 28, lo=np.full(16, 2.0), hi=np.full(16, -2.0))
    ref, _ = O.render_cone(wl.volume, tf, pyr, prm, cam, 28, 28)
    assert cnt[0] > 0 and cnt[1] == cnt[0]
    assert np.allclose(img, ref, atol=1e-13)
    img2, _, cnt2 = O.render_nn(wl.volume, tf, pyr, prm, w, cam, 28, 28)
    assert cnt2[0] == cnt[0] and cnt2[1] == 0
    assert np.abs(img2 - ref).max() > 1e-3  # a random network is not the exact tracer
