"""The headless C++ plugin layer: .raw / .tf1d readers and the RenderingManager driving
a technique through the BaseVolumeRenderer surface (cvr_cli)."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "cpp_volume_rendering_amd" / "cvr_cli"
GOLD = Path(__file__).parent / "golden"


def run(*args, check=True):
    r = subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=300)
    if check and r.returncode != 0:
        raise AssertionError(r.stderr + r.stdout)
    return r


def parse_check(out):
    kv = {}
    for tok in out.replace("\n", " ").split():
        if "=" in tok:
            k, v = tok.split("=", 1)
            kv[k] = v
    return kv


def test_raw_reader_name_encoding(tmp_path):
    vol = np.arange(5 * 4 * 3, dtype=np.uint16).reshape(5, 4, 3) * 1000  # (z, y, x)
    p = tmp_path / "probe.2.3x4x5.raw"
    vol.tofile(p)
    kv = parse_check(run("--volume", p, "--check-inputs").stdout)
    assert "3x4x5" in run("--volume", p, "--check-inputs").stdout
    assert kv["bytes"] == "2"
    assert float(kv["diagonal"]) == pytest.approx(math.sqrt(9 + 16 + 25))
    assert float(kv["sample(0,0,0)"]) == 0.0


def test_raw_reader_rejects_bad_names_and_short_files(tmp_path):
    (tmp_path / "nosizes.raw").write_bytes(b"\0" * 8)
    assert run("--volume", tmp_path / "nosizes.raw", "--check-inputs", check=False).returncode == 1
    (tmp_path / "short.1.4x4x4.raw").write_bytes(b"\0" * 10)
    r = run("--volume", tmp_path / "short.1.4x4x4.raw", "--check-inputs", check=False)
    assert r.returncode == 1 and "short read" in r.stderr


def test_tf1d_reader_bonsai_table():
    kv = parse_check(run("--tf", GOLD / "bonsai_01.tf1d", "--check-inputs").stdout)
    assert kv["max_density"] == "255"
    r, g, b, a = map(float, kv["t[255]"].split(","))
    assert (r, g, b, a) == pytest.approx((0.9, 0.55, 0.0, 0.8), abs=1e-7)
    r, g, b, a = map(float, kv["t[0]"].split(","))
    assert (r, g, b, a) == pytest.approx((0.0, 0.9, 0.0, 0.0), abs=1e-7)
    # GetExtN(0.5) = log(1/(1 - alpha(127.5))): alpha lerps 0.5 -> 0.8 over [40, 255]
    alpha = float(np.float32(0.5 + (0.8 - 0.5) * (127.5 - 40) / 215))
    assert float(kv["extn(0.5)"]) == pytest.approx(-math.log(1 - alpha), rel=1e-5)


def test_tf1d_reader_types_and_errors(tmp_path):
    p = tmp_path / "ext.tf1d"
    p.write_text("linear\n2\n1023 1\n2\n1 0 0 0\n0 0 1 1023\n2\n0.0 0\n2.0 1023\n")
    kv = parse_check(run("--tf", p, "--check-inputs").stdout)
    assert kv["max_density"] == "1023" and kv["extinction_input"] == "1"
    bad = tmp_path / "bad.tf1d"
    bad.write_text("linear\n0\n2\n1 0 0 5\n0 0 1 5\n0\n")
    assert run("--tf", bad, "--check-inputs", check=False).returncode == 1


@pytest.mark.gpu
def test_rendering_manager_frame_matches_c_abi(tmp_path):
    """cvr_cli drives RayCasting1Pass through BaseVolumeRenderer::Init/PrepareRender/Redraw;
    its frame equals the C ABI's render of the same volume, TF and camera."""
    n = 48
    vol = N.synth_blobs((n, n, n), N.U8, n_blobs=6, seed=7)
    raw = tmp_path / f"blobs.1.{n}x{n}x{n}.raw"
    vol.tofile(raw)
    pfm = tmp_path / "frame.pfm"
    csv = tmp_path / "eval.csv"
    run("--technique", "rc1p", "--volume", raw, "--tf", GOLD / "bonsai_01.tf1d", "--size", "64x48",
        "--frames", 2, "--pfm", pfm, "--csv", csv)
    head, data = pfm.read_bytes().split(b"-1.0\n", 1)
    img = np.frombuffer(data, dtype="<f4").reshape(48, 64, 3)
    # the same frame through the C ABI (camera at +z, 2 diagonals away, as InitData sets it)
    r = N.Renderer(N.RC1P)
    r.set_volume(vol)
    r.set_transfer_function([[0, .9, 0, 0], [0, .9, 0, 58], [0, 0, 0, 59], [.7, .35, 0, 60], [.9, .55, 0, 255]],
                            [[0, 0], [0, 35], [.5, 40], [.8, 255]])
    d = float(np.float32(2.0 * math.sqrt(3 * n * n)))
    cam = N.Camera((0.0, 0.0, d), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 64 / 48)
    ref = r.render(cam, 64, 48)
    assert np.abs(img - ref[..., :3]).max() <= 1e-6
    assert ref[..., 3].max() > 0.1
    line = csv.read_text().strip().split(",")
    assert line[0] == "s_1rc" and float(line[1]) > 0


def _multiscale_reference(r, cam, sw, sh, mode, kernel):
    """cvr_render at the mode's frame size, then cvr_resample_frame to the screen."""
    import torch
    sgn = -1 if mode == N.UP_SCALING_RENDER else 1
    fw, fh = N.multiscale_size(sw, sh, 2 * sgn, 2 * sgn)
    frame = torch.zeros((fh, fw, 4), dtype=torch.float32, device="cuda")
    r.render(cam, fw, fh, out=frame)
    dst = torch.zeros((sh, sw, 4), dtype=torch.float32, device="cuda")
    smode = {N.MULTIPLE_RAYS_PER_PIXEL: N.SCALE_MULTISAMPLE, N.DOWN_SCALING_RENDER: N.SCALE_DOWNSCALE,
             N.UP_SCALING_RENDER: N.SCALE_UPSCALE}[mode]
    work = frame.clone() if smode == N.SCALE_UPSCALE else None
    N.resample_frame(frame, dst, smode, kernel, work=work)
    torch.cuda.synchronize()
    return dst.cpu().numpy(), (fw, fh)


@pytest.mark.gpu
@pytest.mark.parametrize("mode,name,kernel,kname", [
    (N.MULTIPLE_RAYS_PER_PIXEL, "multisample", N.FILTER_HAT, "hat"),
    (N.DOWN_SCALING_RENDER, "down", N.FILTER_MITCHELL_NETRAVALI, "mitchell-netravali"),
    (N.DOWN_SCALING_RENDER, "down", N.FILTER_CARDINAL_BSPLINE3, "bspline3"),
    (N.UP_SCALING_RENDER, "up", N.FILTER_CATMULL_ROM, "catmull-rom"),
    (N.UP_SCALING_RENDER, "up", N.FILTER_CARDINAL_OMOMS3, "omoms3"),
])
def test_rendering_manager_multiscale_modes(tmp_path, mode, name, kernel, kname):
    """cvr_cli in each multi-scaling mode (BaseVolumeRenderer::MULTISCALING, dispatched by
    RenderingManager::Display to MultiSampleRedraw / DownScalingRedraw / UpScalingRedraw):
    the screen frame equals the technique's render at twice / half the screen size
    resampled with cvr_resample_frame."""
    n, sw, sh = 40, 66, 50
    vol = N.synth_blobs((n, n, n), N.U8, n_blobs=6, seed=9)
    raw = tmp_path / f"blobs.1.{n}x{n}x{n}.raw"
    vol.tofile(raw)
    pfm = tmp_path / "frame.pfm"
    out = run("--technique", "rc1p", "--volume", raw, "--tf", GOLD / "bonsai_01.tf1d", "--size", f"{sw}x{sh}",
              "--multiscale", name, "--kernel", kname, "--pfm", pfm).stdout
    head, data = pfm.read_bytes().split(b"-1.0\n", 1)
    img = np.frombuffer(data, dtype="<f4").reshape(sh, sw, 3)
    r = N.Renderer(N.RC1P)
    r.set_volume(vol)
    r.set_transfer_function([[0, .9, 0, 0], [0, .9, 0, 58], [0, 0, 0, 59], [.7, .35, 0, 60], [.9, .55, 0, 255]],
                            [[0, 0], [0, 35], [.5, 40], [.8, 255]])
    d = float(np.float32(2.0 * math.sqrt(3 * n * n)))
    cam = N.Camera((0.0, 0.0, d), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, sw / sh)
    ref, (fw, fh) = _multiscale_reference(r, cam, sw, sh, mode, kernel)
    assert f"rendered at {fw}x{fh}" in out
    assert np.abs(img - ref[..., :3]).max() <= 1e-6
    assert ref[..., :3].max() > 0.05
    # the same through the C ABI, to host and to device memory
    host = r.render_multiscale(cam, sw, sh, mode, kernel=kernel)
    assert np.abs(host - ref).max() <= 1e-6
    import torch
    dev = torch.zeros((sh, sw, 4), dtype=torch.float32, device="cuda")
    r.render_multiscale(cam, sw, sh, mode, kernel=kernel, out=dev)
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy(), host)


@pytest.mark.gpu
def test_multiscale_single_ray_and_errors():
    wl = W.plain_blobs(n=24, width=32, height=24, seed=3)
    r = N.Renderer(N.RC1P)
    r.set_volume(wl.volume)
    r.set_transfer_function(wl.rgb_iso, wl.alpha_iso)
    cam = wl.camera()
    assert np.array_equal(r.render_multiscale(cam, 32, 24, N.SINGLE_RAY_PER_PIXEL), r.render(cam, 32, 24))
    for kw in ({"mode": 7}, {"mode": N.DOWN_SCALING_RENDER, "mult": (0, 2)},
               {"mode": N.UP_SCALING_RENDER, "kernel": 9}, {"mode": N.UP_SCALING_RENDER, "mult": (64, 2)}):
        mode = kw.pop("mode")
        with pytest.raises(N.CvrError) as e:
            r.render_multiscale(cam, 32, 24, mode, **kw)
        assert e.value.code == N.ERR_INVALID


def _parse_presets_numpy(path, per_record):
    """Independent parse of a preset file: records of a name line then numbers."""
    lines = [ln.strip() for ln in Path(path).read_text().splitlines() if ln.strip()]
    return lines


def test_camera_state_presets_reference_file():
    """The reference's data/#list_camera_states (a fixture copy): every ARCBALL state, by
    name and by index, with its eye, center and up."""
    f = GOLD / "#list_camera_states"
    lines = _parse_presets_numpy(f, 5)
    names = [lines[i - 1] for i, ln in enumerate(lines) if ln == "ARCBALL"]
    kv = parse_check(run("--camera-states", f, "--camera-state", "Result Bonsai", "--check-inputs").stdout)
    assert int(kv["camera_states"]) == len(names) == 24
    assert kv["type"] == "ARCBALL"
    assert [float(x) for x in kv["eye"].split(",")] == pytest.approx([179.72, 103.536, -292.471])
    assert [float(x) for x in kv["up"].split(",")] == pytest.approx([0, 1, 0])
    out = run("--camera-states", f, "--camera-state", "10", "--check-inputs").stdout
    assert f"camera[10]={names[10]}" in out
    kv = parse_check(out)
    assert [float(x) for x in kv["center"].split(",")] == pytest.approx([19, 32, -7])  # "Result PiggyBank"
    assert [float(x) for x in kv["up"].split(",")] == pytest.approx([0, -1, 0])
    # the tab-indented record
    kv = parse_check(run("--camera-states", f, "--camera-state", "Beechnut", "--check-inputs").stdout)
    assert [float(x) for x in kv["up"].split(",")] == pytest.approx([0, 0, -1])


def test_light_source_presets_reference_file():
    """The reference's data/#list_light_sources: named lists of (position, forward, up,
    right, spot angle) lights; the light keeps forward = -z_axis."""
    f = GOLD / "#list_light_sources"
    lines = _parse_presets_numpy(f, 0)
    names = [lines[i - 1] for i, ln in enumerate(lines) if ln == "1"]
    out = run("--light-sources", f, "--light-source", "Result CT-Knee", "--check-inputs").stdout
    kv = parse_check(out)
    assert int(kv["light_lists"]) == len(names) == 13
    assert kv["lights"] == "1" and float(kv["spot"]) == 5
    assert [float(x) for x in kv["position"].split(",")] == pytest.approx([769.595, -603.05, 592.032])
    fwd = np.array([float(x) for x in kv["forward"].split(",")])
    assert fwd == pytest.approx([0.673312, -0.527603, 0.517963])
    # forward points from the scene towards the light
    pos = np.array([769.595, -603.05, 592.032])
    assert fwd == pytest.approx(pos / np.linalg.norm(pos), abs=1e-5)
    kv = parse_check(run("--light-sources", f, "--light-source", "0", "--check-inputs").stdout)
    assert float(kv["spot"]) == 20


def test_preset_readers_reject_malformed_files(tmp_path):
    bad = tmp_path / "cams"
    bad.write_text("A\nORBIT\n1 2 3\n0 0 0\n0 1 0\n")
    r = run("--camera-states", bad, "--check-inputs", check=False)
    assert r.returncode == 1 and "ARCBALL or FLIGHT" in r.stderr
    bad.write_text("A\nARCBALL\n1 2 3\n0 0\n")
    assert run("--camera-states", bad, "--check-inputs", check=False).returncode == 1
    lights = tmp_path / "lights"
    lights.write_text("L\n1\n0 0 1\n0 0 1\n0 1 0\n")
    assert run("--light-sources", lights, "--check-inputs", check=False).returncode == 1
    lights.write_text("L\n1\n0 0 1\n0 0 1\n0 1 0\n1 0 0\n4\n")
    assert run("--light-sources", lights, "--light-source", "M", "--check-inputs", check=False).returncode == 1
    assert run("--camera-states", tmp_path / "none", "--check-inputs", check=False).returncode == 1


@pytest.mark.gpu
def test_rendering_manager_with_presets(tmp_path):
    """cvr_cli with a camera state and a light-source list of the reference's preset files
    drives the extinction-based renderer (directional light): its frame equals the C
    ABI's render with that camera and light."""
    n = 40
    vol = N.synth_noise((n, n, n), N.U16, octaves=3, seed=5, bands=[0.3, 0.6, 0.85])
    raw = tmp_path / f"noise.2.{n}x{n}x{n}.raw"
    vol.tofile(raw)
    pfm = tmp_path / "frame.pfm"
    run("--technique", "ebs", "--volume", raw, "--tf", GOLD / "bonsai_01.tf1d", "--size", "56x40",
        "--camera-states", GOLD / "#list_camera_states", "--camera-state", "Synthetic Flat 2",
        "--light-sources", GOLD / "#list_light_sources", "--light-source", "Result Backpack", "--pfm", pfm)
    img = np.frombuffer(pfm.read_bytes().split(b"-1.0\n", 1)[1], dtype="<f4").reshape(40, 56, 3)
    r = N.Renderer(N.EBS)
    r.set_volume(vol)
    r.set_transfer_function([[0, .9, 0, 0], [0, .9, 0, 58], [0, 0, 0, 59], [.7, .35, 0, 60], [.9, .55, 0, 255]],
                            [[0, 0], [0, 35], [.5, 40], [.8, 255]])
    r.set_params(shadow_type=1)
    cam = N.Camera((0.785494, 121.176, 123.473), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 56 / 40)
    light = N.Light((154.443, -705.839, -70.8097), (0.212732, -0.972231, -0.097534), (0.020848, -0.0952798, 0.995232),
                    (0.976888, 0.213751, -0.0))
    ref = r.render(cam, 56, 40, light=light)
    assert np.abs(img - ref[..., :3]).max() <= 1e-6
    assert ref[..., 3].max() > 0.1
