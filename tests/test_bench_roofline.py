"""CPU checks of bench.py's roofline helpers: the footprint of every technique's reads
(with stand-in renderers that expose only params() and pyramid()), the memory level a
footprint fits, and the level a gather rate is priced at."""
import types

import numpy as np

import bench
from cpp_volume_rendering_amd import native as N


class FakeRenderer:
    def __init__(self, params, levels=()):
        self._p, self._levels = params, levels

    def params(self):
        return self._p

    def pyramid(self):
        return self._levels


def wl(n=16, w=8, h=4):
    return types.SimpleNamespace(volume=np.zeros((n, n, n), np.uint8), width=w, height=h)


def test_footprint_of_every_technique():
    w = wl()
    base = 16 ** 3 * 8 * 1 + 8 * 4 * 16
    levels = [np.zeros((16, 16, 16), np.float32), np.zeros((8, 8, 8), np.float32)]
    cells = 16 ** 3 + 8 ** 3
    assert bench.read_footprint(N.RC1P, FakeRenderer(None), w, 1, False) == base
    for layout, per_cell in ((0, 64), (1, 8)):
        r = FakeRenderer(types.SimpleNamespace(sat_layout=layout))
        assert bench.read_footprint(N.EBS, r, w, 1, False) == base + 18 ** 3 * per_cell
    for prec, b in ((0, 4), (1, 2), (2, 2)):
        flat = FakeRenderer(types.SimpleNamespace(density_precision=prec), levels)
        assert bench.read_footprint(N.DOS, flat, w, 1, False) == base + cells * 8 * b
        assert bench.read_footprint(N.CONE_TRACE, flat, w, 1, False) == base + cells * 8 * b
        nested = FakeRenderer(types.SimpleNamespace(cone=types.SimpleNamespace(density_precision=prec)), levels)
        assert bench.read_footprint(N.CONE_NN, nested, w, 1, False) == base + cells * 8 * b
    assert bench.read_footprint(N.EBS, FakeRenderer(None), w, 1, True) == base + 32 ** 3 * 8


def test_footprint_levels():
    assert bench.footprint_level(4 << 20) == "l2"
    assert bench.footprint_level((4 << 20) + 1) == "mall"
    assert bench.footprint_level(256 << 20) == "mall"
    assert bench.footprint_level((256 << 20) + 1) == "hbm"


def test_gather_roofline_level_and_step_up():
    peaks = {"l1": 16000.0, "l2": 8000.0, "mall": 1900.0, "hbm": 1560.0}
    r = bench.gather_roofline(2000.0, 1 << 33, peaks, None, 1.0)
    assert r["bound"] == "hbm" and r["peak"] == bench.HBM_PEAK_GBS and r["frac"] == 0.25
    r = bench.gather_roofline(1000.0, 64 << 20, peaks, 5e9, 1.0)
    assert r["bound"] == "mall" and r["peak"] == 1900.0 and r["traffic_gbs"] == 5000.0
    r = bench.gather_roofline(3000.0, 64 << 20, peaks, None, 1.0)  # above the MALL ceiling: L2 serves it
    assert r["bound"] == "l2" and r["footprint_level"] == "mall" and r["frac"] <= 1.0
    r = bench.gather_roofline(9000.0, 1 << 20, peaks, None, 1.0)
    assert r["bound"] == "l1" and r["frac"] <= 1.0


def test_more_than_one_gpu_is_refused():
    """bench.py --gpus N with N > 1 (or a multi-rank launch) exits with a message before it
    touches a device: the build runs on one GPU."""
    import json
    import os
    import subprocess
    import sys
    from pathlib import Path
    root = Path(bench.__file__).resolve().parent
    for args, env in ((["--gpus", "2"], {}), (["--gpus", "1"], {"WORLD_SIZE": "2"})):
        r = subprocess.run([sys.executable, str(root / "bench.py"), *args], capture_output=True, text=True,
                           timeout=120, env={**os.environ, **env})
        assert r.returncode == 2, (r.returncode, r.stderr[-300:])
        assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
