"""Pins the oracle's SAT (reference recurrence) and extinction-based lighting
against numpy prefix sums, brute-force box sums and closed forms (CPU only)."""
import math

import numpy as np
import pytest

import oracle_lib as O


def tf_const(alpha):
    return O.OTF([[1, 1, 1, 0], [1, 1, 1, 255]], [[alpha, 0], [alpha, 255]])


def test_sat_recurrence_equals_cumsum():
    rng = np.random.default_rng(3)
    vol = rng.integers(0, 256, (9, 7, 11), dtype=np.uint8)
    tf = O.OTF([[0, 0, 0, 0], [1, 1, 1, 255]], [[0.0, 0], [0.1, 60], [0.7, 200], [0.9, 255]])
    g = O.sat_input(vol, tf)
    s = O.sat(vol, tf)
    ref = g.cumsum(0).cumsum(1).cumsum(2)
    assert np.allclose(s, ref, rtol=1e-13, atol=1e-12)
    # interior input is GetExtN of the voxel, border zero
    assert np.all(g[0] == 0) and np.all(g[-1] == 0) and np.all(g[:, 0] == 0) and np.all(g[:, :, -1] == 0)
    assert g[1, 1, 1] == pytest.approx(tf.get_ext_n(vol[0, 0, 0] / 255.0))


def test_sat_constant_field_and_box_sums():
    n = (6, 5, 4)  # nx, ny, nz
    vol = np.full((n[2], n[1], n[0]), 77, np.uint8)
    tf = tf_const(0.25)
    e = float(np.float32(-math.log(1.0 - float(np.float32(0.25)))))
    s = O.sat(vol, tf)
    for z in range(1, n[2] + 1):
        for y in range(1, n[1] + 1):
            for x in range(1, n[0] + 1):
                assert s[z, y, x] == pytest.approx(e * x * y * z, rel=1e-13)
    # brute-force box sums from random-valued SAT
    rng = np.random.default_rng(5)
    v2 = rng.integers(0, 256, (n[2], n[1], n[0]), dtype=np.uint8)
    tf2 = O.OTF([[0, 0, 0, 0], [1, 1, 1, 255]], [[0.0, 0], [0.8, 255]])
    g, s2 = O.sat_input(v2, tf2), O.sat(v2, tf2)
    for (x0, x1, y0, y1, z0, z1) in ((0, 3, 1, 4, 0, 2), (2, 6, 0, 5, 1, 4)):
        box = (s2[z1, y1, x1] - s2[z1, y1, x0] - s2[z1, y0, x1] + s2[z1, y0, x0]
               - s2[z0, y1, x1] + s2[z0, y1, x0] + s2[z0, y0, x1] - s2[z0, y0, x0])
        assert box == pytest.approx(g[z0 + 1:z1 + 1, y0 + 1:y1 + 1, x0 + 1:x1 + 1].sum(), rel=1e-12)


def ao_closed_form(e, shells, R=1.0):
    """Uniform extinction e: box sum of half-width r voxels is e*(2r)^3 (the SAT of a
    constant field is multilinear, so GL_LINEAR taps are exact)."""
    S = lambda r: e * (2 * r) ** 3
    t = S(R) / R ** 2
    for i in range(1, shells):
        r1 = R * (i + 1)
        t += (S(r1) - S(R * i)) / r1 ** 2
    rn = R * shells
    return math.exp(-t / rn ** 2)


def test_ebs_ao_uniform_closed_form():
    n = 48
    vol = np.full((n, n, n), 128, np.uint8)
    tf = tf_const(0.05)
    e = float(np.float32(-math.log(1.0 - float(np.float32(0.05)))))
    s = O.sat(vol, tf)
    prm = O.ebs_params(ao_shells=6, apply_shadow=0)
    for pos in ((24.0, 24.0, 24.0), (20.3, 27.7, 22.1)):
        ao, _ = O.ebs_point((n, n, n), (1, 1, 1), s, prm, pos)
        assert ao == pytest.approx(ao_closed_form(e, 6), rel=1e-9)


def test_ebs_shadow_uniform_closed_form():
    """Directional light along +z: every box holds the mean extinction e, so the
    shadow is exp(-e * number_of_boxes); boxes run from z_pos = initial (2) in steps of
    interval (2) while z_pos < max_dist and pos.z + z_pos + interval < G - 0.5."""
    n = 40
    vol = np.full((n, n, n), 128, np.uint8)
    tf = tf_const(0.02)
    e = float(np.float32(-math.log(1.0 - float(np.float32(0.02)))))
    s = O.sat(vol, tf)
    for maxd, pos in ((100.0, (20.2, 19.6, 5.3)), (9.0, (20.0, 20.0, 3.0)), (100.0, (21.0, 17.5, 33.9))):
        prm = O.ebs_params(apply_occlusion=0, shadow_type=1, light_fwd=(0, 0, 1), sdw_max_dist=maxd)
        _, sd = O.ebs_point((n, n, n), (1, 1, 1), s, prm, pos)
        count, zp = 0, 2.0
        while zp < maxd and pos[2] + zp + 2.0 < n - 0.5:
            count += 1
            zp += 2.0
        # the cone's cross-section stays inside the volume here: mean extinction e per box
        assert sd == pytest.approx(math.exp(-e * count), rel=1e-9)


def test_ebs_shadow_axis_selection_symmetry():
    """Light along -x, +y, -z on a volume that is symmetric under the matching axis
    permutation gives the same shadow at the permuted point."""
    n = 30
    rng = np.random.default_rng(7)
    base = rng.integers(0, 256, (n, n, n), dtype=np.uint8)
    tf = O.OTF([[1, 1, 1, 0], [1, 1, 1, 255]], [[0.0, 0], [0.3, 255]])
    # swap x and z axes of the volume (array is (z, y, x))
    sw = np.ascontiguousarray(base.transpose(2, 1, 0))
    s1, s2 = O.sat(base, tf), O.sat(sw, tf)
    p1 = O.ebs_params(apply_occlusion=0, shadow_type=1, light_fwd=(0, 0.2, 1), sdw_max_dist=60)
    p2 = O.ebs_params(apply_occlusion=0, shadow_type=1, light_fwd=(1, 0.2, 0), sdw_max_dist=60)
    _, a = O.ebs_point((n, n, n), (1, 1, 1), s1, p1, (11.3, 14.2, 6.7))
    _, b = O.ebs_point((n, n, n), (1, 1, 1), s2, p2, (6.7, 14.2, 11.3))
    assert a == pytest.approx(b, rel=1e-9)


@pytest.mark.parametrize("light", [("dir", (0.15, 1.0, -0.3)), ("dir", (-0.2, -1.0, 0.25)), ("point", (9.0, 75.0, 12.0)),
                                   ("point", (21.0, -40.0, 5.0))])
def test_ebs_shadow_y_axis_equals_permuted_z_axis(light):
    """A y-dominant light (the ConeYAxis branch) on a volume gives the shadow that the
    z-dominant light of the y <-> z permutation (the ConeZAxis branch, pinned by the
    closed form above) gives on the permuted volume at the permuted point; directional
    and point lights, both signs."""
    n = (26, 30, 22)  # nx, ny, nz: unequal sizes catch a dimension taken from the wrong axis
    rng = np.random.default_rng(17)
    base = rng.integers(0, 256, (n[2], n[1], n[0]), dtype=np.uint8)
    tf = O.OTF([[1, 1, 1, 0], [1, 1, 1, 255]], [[0.0, 0], [0.25, 255]])
    perm = np.ascontiguousarray(base.transpose(1, 0, 2))  # (z, y, x) -> (y, z, x): y and z swapped
    s1, s2 = O.sat(base, tf), O.sat(perm, tf)
    kind, v = light
    sw = (v[0], v[2], v[1])
    if kind == "dir":
        p1 = O.ebs_params(apply_occlusion=0, shadow_type=1, light_fwd=v, sdw_max_dist=80)
        p2 = O.ebs_params(apply_occlusion=0, shadow_type=1, light_fwd=sw, sdw_max_dist=80)
    else:
        # light positions are world coordinates, centred on the box
        p1 = O.ebs_params(apply_occlusion=0, shadow_type=0, light_pos=v, sdw_max_dist=80)
        p2 = O.ebs_params(apply_occlusion=0, shadow_type=0, light_pos=sw, sdw_max_dist=80)
    shadows = []
    for pos in ((11.3, 14.2, 6.7), (4.6, 22.9, 17.2), (20.1, 8.4, 11.0)):
        _, a = O.ebs_point(n, (1, 1, 1), s1, p1, pos)
        _, b = O.ebs_point((n[0], n[2], n[1]), (1, 1, 1), s2, p2, (pos[0], pos[2], pos[1]))
        assert a == pytest.approx(b, rel=1e-9), pos
        shadows.append(a)
    assert min(shadows) < 0.95  # the medium casts a shadow at some of the points
