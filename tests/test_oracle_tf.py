"""Pins the oracle's transfer function and 1D texture sampling against closed forms
and the reference's own TF file (data/tf1dcp/bonsai_01.tf1d, kept as a fixture)."""
import math
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import OTF, lib

GOLD = Path(__file__).parent / "golden"


def parse_tf1d(path):
    """Format of transferfunction1d.h:1-19 (type 0: max density 255)."""
    tok = path.read_text().split("\n", 1)[1].split()
    it = iter(tok)
    typ = int(next(it))
    md, ext = 255, 0
    if typ in (1, 2):
        md = int(next(it))
    if typ == 2:
        ext = int(next(it))
    n = int(next(it))
    rgb = [[float(next(it)), float(next(it)), float(next(it)), float(next(it))] for _ in range(n)]
    m = int(next(it))
    a = [[float(next(it)), float(next(it))] for _ in range(m)]
    return np.array(rgb), np.array(a), md, ext


def test_bonsai_control_points_exact():
    rgb, a, md, ext = parse_tf1d(GOLD / "bonsai_01.tf1d")
    tf = OTF(rgb, a, md, ext)
    tab, rgbt, rgba = tf.tables()
    assert tab.shape == (256, 4)
    # colours at control points are the float-rounded control values
    for r, g, b, iso in rgb:
        assert tab[int(iso), 0] == float(np.float32(r))
        assert tab[int(iso), 1] == float(np.float32(g))
        assert tab[int(iso), 2] == float(np.float32(b))
    for al, iso in a:
        assert tab[int(iso), 3] == float(np.float32(al))
    # piecewise linear: second differences vanish inside each segment
    for col, pts in ((1, rgb[:, 3]), (0, rgb[:, 3]), (3, a[:, 1])):
        for i0, i1 in zip(pts[:-1], pts[1:]):
            seg = tab[int(i0):int(i1) + 1, col]
            if len(seg) >= 3:
                assert np.abs(np.diff(seg, 2)).max() < 1e-12
    # hand values: alpha at 37 = 0.5 * 2/5, green between 0 and 58 is 0.9f
    assert tab[37, 3] == pytest.approx(0.2, abs=1e-7)
    assert tab[30, 1] == pytest.approx(0.9, abs=1e-7)


def test_rgbt_is_extinction_and_rgba_is_opacity():
    rgb, a, md, ext = parse_tf1d(GOLD / "bonsai_01.tf1d")
    tab, rgbt, rgba = OTF(rgb, a, md, ext).tables()
    for i in (0, 36, 37, 40, 100, 255):
        op = float(np.float32(tab[i, 3]))
        assert rgbt[i, 3] == pytest.approx(-math.log(1.0 - op), rel=1e-6, abs=1e-7)
        assert rgba[i, 3] == pytest.approx(op, abs=1e-7)


def test_get_and_get_ext_n():
    rgb, a, md, ext = parse_tf1d(GOLD / "bonsai_01.tf1d")
    tf = OTF(rgb, a, md, ext)
    tab = tf.tables()[0]
    # integer positions return the table, midpoints the average, out of range zero
    assert np.allclose(tf.get(100.0), tab[100], atol=1e-7)
    assert np.allclose(tf.get(100.5), 0.5 * (tab[100] + tab[101]), atol=1e-7)
    assert np.allclose(tf.get(255.0), tab[255], atol=1e-7)
    assert np.all(tf.get(-1.0) == 0) and np.all(tf.get(256.0) == 0)
    # normalised: v in [0,1] maps to v*255
    assert tf.get_ext_n(1.0) == pytest.approx(-math.log(1 - 0.8), rel=1e-6)
    assert tf.get_ext_n(37 / 255) == pytest.approx(-math.log(1 - 0.2), rel=1e-5)


def test_duplicate_isovalues_rejected():
    with pytest.raises(ValueError):
        OTF([[0, 0, 0, 0], [1, 1, 1, 10], [1, 1, 1, 10]], [[0, 0], [1, 255]])


def test_tex1d_texel_centres():
    T = 256
    rng = np.random.default_rng(0)
    table = rng.uniform(0, 1, (T, 4)).astype(np.float32)
    out = np.zeros(4)
    L = lib()
    for i in (0, 1, 17, 254, 255):
        L.orc_tex1d(table.ctypes.data, T, (i + 0.5) / T, out.ctypes.data)
        assert np.allclose(out, table[i], atol=1e-12)
    L.orc_tex1d(table.ctypes.data, T, 0.0, out.ctypes.data)
    assert np.allclose(out, table[0])
    L.orc_tex1d(table.ctypes.data, T, 1.0, out.ctypes.data)
    assert np.allclose(out, table[T - 1])
    L.orc_tex1d(table.ctypes.data, T, 100 / T, out.ctypes.data)
    assert np.allclose(out, 0.5 * (table[99].astype(np.float64) + table[100]), atol=1e-12)
