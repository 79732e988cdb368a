"""The learned occlusion/shadow surrogate (CVR_CONE_NN): bf16-MFMA inference and
training kernels against the FP64 MLP of the oracle, and the trained renderer
against the exact cone tracer it replaces."""
import numpy as np
import pytest

import oracle_lib as O
from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def random_net(seed, scale=0.25):
    rng = np.random.default_rng(seed)
    P = N.lib().cvr_nn_param_count()
    w = rng.normal(0, scale, P).astype(np.float32)
    return w


def ref_forward_bf16(w, X):
    """FP64 forward with the bf16 rounding points of the GPU: weights and layer inputs
    rounded to bf16, products accumulated exactly, biases in fp32."""
    lay = [(64, 16), (64, 64), (64, 64), (16, 64)]
    o = 0
    a = bf16(X)
    for l, (out, inp) in enumerate(lay):
        Wm = bf16(w[o:o + out * inp]).reshape(out, inp)
        o += out * inp
        b = w[o:o + out].astype(np.float64)
        o += out
        z = a @ Wm.T + b
        a = bf16(np.maximum(z, 0)) if l < 3 else z
    return a[:, :2]


@pytest.fixture(scope="module")
def nn_handle():
    wl = W.cone_blobs(n=32, width=8, height=8)
    r = N.Renderer(N.CONE_NN)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    return r


@pytest.mark.parametrize("path", [0, 1])
def test_inference_matches_fp64(nn_handle, path):
    w = random_net(1)
    nn_handle.nn_set_weights(w)
    X = np.random.default_rng(2).uniform(0, 1, (1000, 16)).astype(np.float32)
    Y = nn_handle.nn_eval(X, path)
    ref = ref_forward_bf16(w, X)
    full = O.mlp_forward(w.astype(np.float64), X.astype(np.float64))[:, :2]
    print(np.abs(Y - ref).max(), np.abs(Y - full).max())
    assert np.abs(Y - ref).max() <= 2e-3 * (1 + np.abs(ref).max())
    assert np.abs(Y - full).max() <= 3e-2 * (1 + np.abs(full).max())


def test_register_path_equals_training_path(nn_handle):
    """The renderer's register path (16x16x32 MFMA chunks, hardware bf16 rounding) and the
    training kernels' path (16x16x16 MFMA from LDS) round the same bf16 activations; fp32
    sums in another order can move an activation across a bf16 rounding boundary, so
    they agree to bf16 resolution."""
    nn_handle.nn_set_weights(random_net(3))
    X = np.random.default_rng(4).uniform(0, 1, (777, 16)).astype(np.float32)
    a, b = nn_handle.nn_eval(X, 0), nn_handle.nn_eval(X, 1)
    print(np.abs(a - b).max())
    assert np.abs(a - b).max() <= 2e-3 * (1 + np.abs(a).max())


@pytest.mark.parametrize("n", [1024, 1000, 130])  # whole 64-sample chunks, and a last chunk of 40 or 2
def test_gradient_matches_fp64(nn_handle, n):
    w = random_net(5, 0.2)
    nn_handle.nn_set_weights(w)
    rng = np.random.default_rng(6)
    X = rng.uniform(0, 1, (n, 16)).astype(np.float32)
    T = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    loss, g = nn_handle.nn_gradient(X, T)
    rl, rg = O.mlp_loss_grad(w.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert loss == pytest.approx(rl, rel=3e-2)
    offs = [0, 1024, 1088, 5184, 5248, 9344, 9408, 10432, 10448]
    for a, b in zip(offs[:-1], offs[1:]):
        ga, ra = g[a:b].astype(np.float64), rg[a:b]
        rel = np.linalg.norm(ga - ra) / max(np.linalg.norm(ra), 1e-12)
        cos = ga @ ra / max(np.linalg.norm(ga) * np.linalg.norm(ra), 1e-30)
        print(a, b, rel, cos)
        assert rel <= 5e-2 and cos >= 0.998, (a, b, rel, cos)


def train_small(n=64, n_train=1 << 19, epochs=4, seed=21):
    wl = W.cone_blobs(n=n, width=96, height=96, seed=seed)
    r = N.Renderer(N.CONE_NN)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    p = r.params()
    p.n_train, p.n_val, p.epochs, p.batch = n_train, 1 << 14, epochs, 4096
    p.cone.n_levels = 6
    r.set_params(p)
    r.nn_train()
    return wl, r, p


def test_training_beats_the_network_free_baseline():
    wl, r, p = train_small()
    rep = r.nn_report()
    print(rep.train_loss, rep.val_rmse_occ, rep.val_rmse_sdw, rep.base_rmse_occ, rep.base_rmse_sdw)
    assert rep.steps == (p.n_train // p.batch) * p.epochs
    assert rep.val_rmse_occ < 0.05 and rep.val_rmse_sdw < 0.05
    # what the network gains over the network-free estimate from the same lookups
    assert rep.val_rmse_occ < 0.6 * rep.base_rmse_occ
    assert rep.val_rmse_sdw < 0.8 * rep.base_rmse_sdw
    assert rep.exact_fallback == 0


def test_surrogate_image_close_to_exact_tracer():
    wl, r, p = train_small()
    light = wl.light()
    cam = wl.camera()
    r.collect_stats(True)
    img = r.render(cam, 96, 96, light=light)
    st = r.stats()
    ct = N.Renderer(N.CONE_TRACE)
    ct.set_volume(wl.volume)
    ct.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    ct.set_params(p.cone)
    ref = ct.render(cam, 96, 96, light=light)
    rmse = float(np.sqrt(np.mean((img - ref) ** 2)))
    print("rmse", rmse, "fallback", st.fallback_samples, "shaded", st.shaded_samples)
    assert r.nn_report().exact_fallback == 0  # the network rendered, not the exact tracer
    assert rmse <= 0.01
    assert st.shaded_samples > 0 and st.fallback_samples <= 0.05 * st.shaded_samples
    # forcing the exact tracer reproduces the cone-traced renderer (no retraining)
    p.force_exact = 1
    r.set_params(p)
    exact = r.render(cam, 96, 96, light=light)
    assert np.array_equal(exact, ref)


def test_out_of_range_samples_take_the_exact_tracer():
    """The renderer's first pass uses the network alone and marks the tiles where a
    sample's inputs leave the training range (+- ood_margin); a second pass renders the
    marked tiles again with the exact tracer for those samples and clears the marks."""
    wl, r, p = train_small()
    cam, light = wl.camera(), wl.light()
    r.collect_stats(True)
    base = r.render(cam, 96, 96, light=light)
    st0 = r.stats()
    assert st0.fallback_samples == 0
    ct = N.Renderer(N.CONE_TRACE)
    ct.set_volume(wl.volume)
    ct.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    ct.set_params(p.cone)
    ref = ct.render(cam, 96, 96, light=light)
    # an empty range: every shaded sample is out of it, every tile is rendered again and
    # the frame is the exact tracer's
    p.ood_margin = -2.0
    r.set_params(p)
    allx = r.render(cam, 96, 96, light=light)
    st = r.stats()
    print("all out of range: max diff", np.abs(allx - ref).max(), st.fallback_samples, st.shaded_samples)
    assert st.fallback_samples == st.shaded_samples == st0.shaded_samples
    assert st.samples == st0.samples
    assert np.abs(allx - ref).max() <= 2e-5  # the two renderers round wgt * src * m in another order
    # a narrowed range: some tiles marked, the others the network's
    p.ood_margin = -0.02
    r.set_params(p)
    part = r.render(cam, 96, 96, light=light)
    st = r.stats()
    print("narrowed range: fallback", st.fallback_samples, "of", st.shaded_samples)
    assert 0 < st.fallback_samples < st.shaded_samples
    changed = np.any(part != base, axis=-1)
    assert changed.any() and not changed.all()
    assert np.array_equal(r.render(cam, 96, 96, light=light), part)  # marks were cleared in between
    # the range restored: no marks left over, the first frame again
    p.ood_margin = 0.02
    r.set_params(p)
    again = r.render(cam, 96, 96, light=light)
    assert r.stats().fallback_samples == 0
    assert np.array_equal(again, base)


def test_checkpoint_restores_the_renderer_exactly():
    wl, r, p = train_small(n=32, n_train=1 << 15, epochs=2)
    blob = r.nn_save()
    assert blob[:8] == b"CVRNNCK1"
    cam, light = wl.camera(), wl.light()
    a = r.render(cam, 48, 48, light=light)
    r2 = N.Renderer(N.CONE_NN)
    r2.set_volume(wl.volume)
    r2.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    r2.set_params(p)
    r2.nn_load(blob)
    b = r2.render(cam, 48, 48, light=light)
    assert np.array_equal(a, b)
    assert np.array_equal(r2.nn_weights(), r.nn_weights())
    with pytest.raises(N.CvrError):
        r2.nn_load(b"X" * len(blob))


# ---------------------------------------------------------------------------- frame parity
import surrogate_oracle as SO  # noqa: E402


def assert_frame_parity(img, ref, m, tol, stride, max_over_frac):
    s = O.compare(img, ref, m, tol, stride)
    print(s)
    assert s["max_abs"] <= tol or s["n_over_tol_nonstraddle"] <= max_over_frac * s["n_pixels"], s
    assert s["n_over_tol"] <= max_over_frac * s["n_pixels"] + 1, s
    return s


def test_surrogate_frame_matches_fp64_oracle():
    """The surrogate renderer's whole frame against the FP64 surrogate frame on the same
    pyramid, weights and input range: with the network's bf16 rounding points in the
    oracle the frames agree within 2e-3 per channel; against the plain FP64 network (no
    bf16 rounding) within 1e-2 (the bf16 network's own error)."""
    wl, r, p = train_small()
    cam, light = wl.camera(), wl.light()
    r.collect_stats(True)
    img = r.render(cam, 96, 96, light=light)
    st = r.stats()
    ref, m, cnt = SO.nn_oracle_frame(wl, r, cam, light, 96, 96)
    print("shaded", st.shaded_samples, cnt[0], "exact", st.fallback_samples, cnt[1])
    assert ref[..., 3].max() > 0.2
    assert abs(int(cnt[0]) - int(st.shaded_samples)) <= 0.002 * st.shaded_samples
    assert_frame_parity(img, ref, m, 2e-3, 1, 0.005)
    ref64, m64, _ = SO.nn_oracle_frame(wl, r, cam, light, 96, 96, bf16=False)
    assert_frame_parity(img, ref64, m64, 1e-2, 1, 0.005)


def test_surrogate_frame_camera_inside_matches_fp64_oracle():
    """A camera inside the volume box (rays start at the eye, not at the box entry) and a
    wide field of view: the surrogate frame against the FP64 surrogate frame."""
    wl, r, p = train_small()
    cam = N.Camera((3.0, -2.0, 5.0), (0.0, 0.0, -40.0), (0.0, 1.0, 0.0), 60.0, 1.0)
    light = wl.light()
    img = r.render(cam, 64, 64, light=light)
    ref, m, _ = SO.nn_oracle_frame(wl, r, cam, light, 64, 64)
    assert ref[..., 3].max() > 0.2
    assert_frame_parity(img, ref, m, 2e-3, 1, 0.005)


@pytest.mark.parametrize("prec", [1, 2])
def test_surrogate_frame_16bit_pyramid_matches_fp64_oracle(prec):
    """fp16 / bf16 pyramid storage: the oracle reads the same rounded values."""
    wl = W.cone_blobs(n=48, width=64, height=64, seed=22)
    r = N.Renderer(N.CONE_NN)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    p = r.params()
    p.n_train, p.n_val, p.epochs, p.batch = 1 << 17, 1 << 13, 2, 4096
    p.cone.n_levels, p.cone.density_precision = 6, prec
    r.set_params(p)
    r.nn_train()
    cam, light = wl.camera(), wl.light()
    img = r.render(cam, 64, 64, light=light)
    ref, m, _ = SO.nn_oracle_frame(wl, r, cam, light, 64, 64)
    assert_frame_parity(img, ref, m, 2e-3, 1, 0.005)


@pytest.fixture(scope="module")
def cfg4():
    """The surrogate config at full size: 512^3 u8 (64 blobs + 4-octave noise, seed 4),
    1024^2, 7-level pyramid, network trained on 2^22 exact-tracer samples."""
    wl = W.surrogate_blobs_noise()
    r = N.Renderer(N.CONE_NN)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    r.prepare()
    return wl, r


def test_cfg4_pyramid_sampled_voxels(cfg4):
    wl, r = cfg4
    levels = r.pyramid()
    assert [a.shape for a in levels] == [(512 >> i,) * 3 for i in range(7)]
    SO.check_pyramid_sparse(wl, levels, (512, 512, 512), np.random.default_rng(7))


def test_cfg4_surrogate_frame_matches_fp64_oracle(cfg4):
    """512^3 at 1024^2: the HIP surrogate frame against the FP64 surrogate frame on every
    32nd pixel (x and y), and the sample counts of the march."""
    wl, r = cfg4
    assert r.nn_report().exact_fallback == 0
    cam, light = wl.camera(), wl.light()
    r.collect_stats(True)
    img = r.render(cam, 1024, 1024, light=light)
    st = r.stats()
    print("samples", st.samples, "shaded", st.shaded_samples, "exact", st.fallback_samples)
    assert st.fallback_samples <= 0.01 * st.shaded_samples
    ref, m, cnt = SO.nn_oracle_frame(wl, r, cam, light, 1024, 1024, stride=32)
    assert ref[..., 3].max() > 0.2
    assert_frame_parity(img, ref, m, 2e-3, 32, 0.005)


def test_cfg4_fp32_pyramid_wide_offsets_match_fp64_oracle(cfg4):
    """fp32 pyramid storage at 512^3: level 0's corner blocks span 2^27 cells x 32 bytes
    = 4 GiB, so the feature reads take the 64-bit record offsets (the fp16 default takes
    the 32-bit ones); the frame against the FP64 surrogate frame on every 64th pixel."""
    wl, _ = cfg4
    assert 512 ** 3 * 32 > 0xFFFFFFFF
    r = N.Renderer(N.CONE_NN)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    p = r.params()
    p.cone.density_precision = 0
    p.n_train, p.epochs = 1 << 20, 2
    p.rmse_gate = 1.0  # the network renders whatever its validation error
    r.set_params(p)
    r.prepare()
    cam, light = wl.camera(), wl.light()
    img = r.render(cam, 1024, 1024, light=light)
    ref, m, _ = SO.nn_oracle_frame(wl, r, cam, light, 1024, 1024, stride=64)
    assert ref[..., 3].max() > 0.2
    assert_frame_parity(img, ref, m, 2e-3, 64, 0.005)


def test_cfg4_image_error_against_exact_tracer(cfg4):
    """The acceptance figure of the surrogate: image RMSE <= 0.01 against the exact cone
    tracer at 512^3 and 1024^2, over the whole frame and over RGB of covered pixels; and
    what the trained network gains over the network-free estimate of the same inputs
    (the same renderer with a zero network) on that frame."""
    wl, r = cfg4
    cam, light = wl.camera(), wl.light()
    img = r.render(cam, 1024, 1024, light=light)
    ct = N.Renderer(N.CONE_TRACE)
    ct.set_volume(wl.volume)
    ct.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    ct.set_params(r.params().cone)
    ref = ct.render(cam, 1024, 1024, light=light)
    whole, rgb, ncov = SO.image_rmse(img, ref)
    print("image RMSE vs exact: whole frame", whole, "RGB over", ncov, "covered pixels", rgb)
    assert whole <= 0.01 and rgb <= 0.01
    assert ncov > 0.2 * 1024 * 1024
    blob = r.nn_save()
    try:
        r.nn_set_weights(np.zeros(N.lib().cvr_nn_param_count(), np.float32))
        base = r.render(cam, 1024, 1024, light=light)
    finally:
        r.nn_load(blob)
    bw, brgb, _ = SO.image_rmse(base, ref)
    print("network-free estimate: whole", bw, "RGB covered", brgb)
    assert rgb < 0.7 * brgb
    # and on the validation samples of each output, shadows included
    rep = r.nn_report()
    print("validation RMSE occ/sdw", rep.val_rmse_occ, rep.val_rmse_sdw, "network-free", rep.base_rmse_occ,
          rep.base_rmse_sdw)
    assert rep.val_rmse_occ < 0.5 * rep.base_rmse_occ and rep.val_rmse_sdw < 0.9 * rep.base_rmse_sdw
    assert np.array_equal(r.render(cam, 1024, 1024, light=light), img)  # the checkpoint restored the network
