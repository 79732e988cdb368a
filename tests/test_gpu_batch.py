"""Batched surrogate rendering (cvr_batch_*): one persistent graph-captured launch over
uneven volumes, views and lights gives, image by image, what the single-frame renderer
gives for the same handle, camera and light (the same device code inlined into another
kernel: equal up to fp32 rounding, within 1e-4), and the same bits on every graph
replay, for every work shape and after autotuning."""
import numpy as np
import pytest
import torch

from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W

pytestmark = pytest.mark.gpu
WD, HT = 64, 48
EAGER_TOL = 1e-4


def surrogate(vol, rgb, a, seed):
    r = N.Renderer(N.CONE_NN)
    r.set_volume(vol)
    r.set_transfer_function(rgb, a)
    p = r.params()
    p.n_train, p.n_val, p.epochs, p.batch, p.seed = 1 << 16, 1 << 12, 2, 4096, seed
    p.rmse_gate = 1.0  # the network renders (the gate is tested elsewhere)
    r.set_params(p)
    r.prepare()
    return r


@pytest.fixture(scope="module")
def scene():
    vols = [
        N.synth_blobs((40, 40, 40), N.U8, n_blobs=12, seed=31),
        N.synth_noise((48, 32, 24), N.U16, octaves=3, seed=32, bands=[0.3, 0.6, 0.85]),
        N.synth_blobs((32, 36, 28), N.U8, n_blobs=8, seed=33, noise_weight=0.25),
    ]
    rs = []
    for i, v in enumerate(vols):
        rgb, a = W.seeded_tf(40 + i, 4 + 2 * i)
        rs.append(surrogate(v, rgb, a, 50 + i))
    jobs = []
    for i, v in enumerate(vols):
        nz, ny, nx = v.shape
        grid = (float(nx), float(ny), float(nz))
        for az in (10.0, 130.0, 250.0):
            for ld in ((-0.45, 0.75, 0.5), (0.6, 0.3, -0.7)):
                jobs.append((i, W.orbit_camera(grid, WD, HT, az, 25.0), W.light_from_direction(ld, grid)))
    ref = np.stack([rs[v].render(cam, WD, HT, light=l) for v, cam, l in jobs])
    return rs, jobs, ref


def batch_reference(b, n):
    """The batch's images at the default work shape (rendered once per batch)."""
    t = b.tuning()
    b.set_tuning(tile_w=8, shade_batch=0.5, queues=8, workgroups_per_cu=2, span=1)
    out = torch.zeros((n, HT, WD, 4), device="cuda")
    b.render(out)
    torch.cuda.synchronize()
    b.set_tuning(tile_w=t.tile_w, shade_batch=t.shade_batch, queues=t.queues, workgroups_per_cu=t.workgroups_per_cu,
                 span=t.span)
    return out.cpu().numpy()


def make_batch(rs, jobs):
    b = N.Batch(WD, HT)
    for r in rs:
        b.add_volume(r)
    b.set_jobs(jobs)
    return b


def test_graph_replays_equal_single_frames(scene):
    rs, jobs, ref = scene
    assert ref[..., 3].max() > 0.2
    b = make_batch(rs, jobs)
    out = torch.zeros((len(jobs), HT, WD, 4), device="cuda")
    first = None
    for rep in range(3):
        out.zero_()
        b.render(out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        err = float(np.abs(got - ref).max())
        print("replay", rep, "max-abs vs single frames", err)
        assert err <= EAGER_TOL, (rep, err)
        if first is None:
            first = got
        assert np.array_equal(got, first), rep
        # eager work between replays: a single-frame render and a host copy
        rs[rep % len(rs)].render(jobs[0][1], WD, HT, light=jobs[0][2])
    st = b.stats()
    assert st.shaded_samples > 0 and st.fallback_samples <= st.shaded_samples


@pytest.mark.parametrize("tile_w,shade,queues,wg,span", [(16, 0.0, 1, 1, 1), (4, 1.0, 8, 3, 8), (32, 0.75, 8, 2, 2)])
def test_work_shapes_do_not_change_pixels(scene, tile_w, shade, queues, wg, span):
    rs, jobs, ref = scene
    b = make_batch(rs, jobs)
    b.set_tuning(tile_w=tile_w, shade_batch=shade, queues=queues, workgroups_per_cu=wg, span=span)
    out = torch.zeros((len(jobs), HT, WD, 4), device="cuda")
    b.render(out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert float(np.abs(got - ref).max()) <= EAGER_TOL
    assert np.array_equal(got, batch_reference(b, len(jobs)))


def test_autotune_then_render(scene):
    rs, jobs, ref = scene
    b = make_batch(rs, jobs)
    out = torch.zeros((len(jobs), HT, WD, 4), device="cuda")
    t = b.autotune(out, stride=2)
    assert t.ms > 0 and t.tile_w in (4, 8, 16, 32)
    out.zero_()
    b.render(out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert float(np.abs(got - ref).max()) <= EAGER_TOL
    assert np.array_equal(got, batch_reference(b, len(jobs)))


def default_step(scale=(1.0, 1.0, 1.0)):
    """The default march step as the library forms it (rc1prenderer.cpp:62-63):
    (0.5f / sqrtf(3)) * |voxel scale|, evaluated in double and stored as float."""
    sx, sy, sz = scale
    return np.float32(float(np.float32(0.5) / np.sqrt(np.float32(3.0))) * np.sqrt(sx * sx + sy * sy + sz * sz))


def frames_at_step(rs, jobs, scale):
    """Single-frame renders with every member's step multiplied by `scale` in float32
    (what a batch at step_scale `scale` computes; the scene's volumes have unit voxels);
    the members' steps are restored."""
    saved = [r.params().cone.step_size for r in rs]
    try:
        for r in rs:
            p = r.params()
            p.cone.step_size = float(default_step() * np.float32(scale))
            r.set_params(p)
        return np.stack([rs[v].render(cam, WD, HT, light=l) for v, cam, l in jobs])
    finally:
        for r, st in zip(rs, saved):
            p = r.params()
            p.cone.step_size = st
            r.set_params(p)


def test_step_scale_equals_single_frames_at_the_scaled_step(scene):
    """step_scale multiplies every member's march step: the batch's images equal
    single-frame renders at the multiplied step, and changing the step does not retrain
    the members' networks."""
    rs, jobs, ref = scene
    before = [r.nn_report().steps for r in rs]
    b = make_batch(rs, jobs)
    b.set_tuning(step_scale=2.0)
    out = torch.zeros((len(jobs), HT, WD, 4), device="cuda")
    b.render(out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = frames_at_step(rs, jobs, 2.0)
    err = float(np.abs(got - want).max())
    print("step_scale 2: max-abs vs single frames at twice the step", err)
    assert err <= EAGER_TOL
    assert float(np.abs(got - ref).max()) > 10 * EAGER_TOL  # the step did change the pixels
    assert [r.nn_report().steps for r in rs] == before


def test_autotune_searches_the_step_schedule(scene):
    """With max_step_rmse set the tuner searches the step first: the step_scale it keeps
    adds at most that image RMSE (RGB over covered pixels) against step_scale 1, as the
    tuner reports it and as measured here over every job, and a render afterwards gives
    the single-frame images at that step."""
    from surrogate_oracle import image_rmse
    rs, jobs, ref = scene
    b = make_batch(rs, jobs)
    out = torch.zeros((len(jobs), HT, WD, 4), device="cuda")
    gate = 0.01
    t = b.autotune(out, stride=1, max_step_rmse=gate)
    print("tuned step_scale", t.step_scale, "rmse", t.step_rmse, "ms", t.ms)
    assert t.step_scale in (1.0, 1.25, 1.5, 2.0, 2.5, 3.0) and 0.0 <= t.step_rmse <= gate
    assert b.tuning().step_scale == t.step_scale
    out.zero_()
    b.render(out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    _, rgb, ncov = image_rmse(got, ref)
    assert ncov > 0 and rgb <= gate + 1e-4
    assert abs(rgb - t.step_rmse) <= 1e-4
    want = frames_at_step(rs, jobs, t.step_scale)
    assert float(np.abs(got - want).max()) <= EAGER_TOL


def test_batch_errors(scene):
    rs, jobs, _ = scene
    b = N.Batch(WD, HT)
    plain = N.Renderer(N.RC1P)
    with pytest.raises(N.CvrError):
        b.add_volume(plain)
    b.add_volume(rs[0])
    with pytest.raises(N.CvrError):
        b.set_jobs([(1, jobs[0][1], jobs[0][2])])
    with pytest.raises(N.CvrError):
        b.set_tuning(tile_w=12)
    with pytest.raises(N.CvrError):
        b.set_tuning(step_scale=8.0)
    with pytest.raises(N.CvrError):
        b.set_tuning(max_step_rmse=-0.5)
    with pytest.raises(ValueError):
        b.set_tuning(step=2.0)
    out = torch.zeros((1, HT, WD, 4), device="cuda")
    with pytest.raises(N.CvrError):
        b.render(out)  # no jobs yet


def test_member_change_rebuilds_the_batch():
    """Changing a member handle between two batch renders (a new volume of another size,
    which frees and reallocates its device buffers and retrains its network; then a new
    transfer function) must not replay the old graph against freed memory: the batch
    picks up the change and still equals the single-frame renderer."""
    vols = [N.synth_blobs((32, 32, 32), N.U8, n_blobs=6, seed=61), N.synth_blobs((24, 28, 20), N.U8, n_blobs=5, seed=62)]
    rs = []
    for i, v in enumerate(vols):
        rgb, a = W.seeded_tf(70 + i, 4)
        rs.append(surrogate(v, rgb, a, 80 + i))
    grid = (32.0, 32.0, 32.0)
    jobs = [(v, W.orbit_camera(grid, WD, HT, az, 25.0), W.light_from_direction((-0.45, 0.75, 0.5), grid))
            for v in range(2) for az in (0.0, 120.0)]
    b = make_batch(rs, jobs)
    out = torch.zeros((len(jobs), HT, WD, 4), device="cuda")
    b.render(out)
    torch.cuda.synchronize()
    # a new volume for member 0: its corner blocks, pyramid and network are rebuilt
    rs[0].set_volume(N.synth_blobs((40, 36, 44), N.U8, n_blobs=9, seed=63))
    rs[0].prepare()
    for step in range(2):
        out.zero_()
        b.render(out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        ref = np.stack([rs[v].render(cam, WD, HT, light=l) for v, cam, l in jobs])
        assert float(np.abs(got - ref).max()) <= EAGER_TOL, step
        assert ref[..., 3].max() > 0.05
        rgb, a = W.seeded_tf(90, 6)
        rs[1].set_transfer_function(rgb, a)  # a new transfer function (and network) for the second round


# ---------------------------------------------------------------------------- the full scene
import oracle_lib as O  # noqa: E402
import surrogate_oracle as SO  # noqa: E402

FULL = 512


@pytest.fixture(scope="module")
def full_scene():
    """The batched config at full size: 128^3 u8, 192x256x160 u16, 384^3 u8 and 512^3 u16
    volumes, each with its own seeded TF and a surrogate trained on 2^22 exact-tracer
    samples; 16 views x 4 lights per volume, 256 images of 512^2 in one launch."""
    wls, jobs = W.batch_scene()
    rs = []
    for i, wl in enumerate(wls):
        r = N.Renderer(N.CONE_NN)
        r.set_volume(wl.volume, wl.scale)
        r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
        r.prepare()
        rs.append(r)
    b = N.Batch(FULL, FULL)
    for r in rs:
        b.add_volume(r)
    b.set_jobs(jobs)
    out = torch.zeros((len(jobs), FULL, FULL, 4), device="cuda")
    b.render(out)
    torch.cuda.synchronize()
    return wls, rs, jobs, out


def test_full_batch_equals_single_frames(full_scene):
    """Every 8th of the 256 images against the single-frame renderer of its handle."""
    wls, rs, jobs, out = full_scene
    assert len(jobs) == 256 and out.shape == (256, FULL, FULL, 4)
    worst = 0.0
    for j in range(0, len(jobs), 8):
        v, cam, light = jobs[j]
        ref = rs[v].render(cam, FULL, FULL, light=light)
        got = out[j].cpu().numpy()
        err = float(np.abs(got - ref).max())
        worst = max(worst, err)
        assert err <= EAGER_TOL, (j, err)
    print("max-abs vs single frames over 32 images", worst)
    assert out[..., 3].amax().item() > 0.2


def test_full_batch_one_job_per_volume_matches_fp64_oracle(full_scene):
    """One image per volume (a different view and light each) against the FP64 surrogate
    frame of its handle on every 16th pixel."""
    wls, rs, jobs, out = full_scene
    for v in range(len(wls)):
        j = [i for i, jb in enumerate(jobs) if jb[0] == v][5 + 3 * v]
        _, cam, light = jobs[j]
        img = out[j].cpu().numpy()
        ref, m, cnt = SO.nn_oracle_frame(wls[v], rs[v], cam, light, FULL, FULL, stride=16)
        s = O.compare(img, ref, m, 2e-3, 16)
        print(wls[v].name, "job", j, s, "oracle exact-tracer samples", cnt[1], "of", cnt[0])
        assert ref[..., 3].max() > 0.1
        assert s["max_abs"] <= 2e-3 or s["n_over_tol_nonstraddle"] <= 0.005 * s["n_pixels"], s
        assert s["n_over_tol"] <= 0.005 * s["n_pixels"] + 1, s
