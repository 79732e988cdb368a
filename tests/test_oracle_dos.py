"""Pins the oracle's directional-occlusion section tables and integration, and its
ray-sampled cone ground truth, against closed forms and invariants (CPU only)."""
import math

import numpy as np
import pytest

import oracle_lib as O

DIV3 = 1.0 + 2.0 / math.sqrt(3.0)
DIV7 = 3.01


def spec(half=20.0, packing=1, weight=0.35, init=3.0, covered=110.0, d_sigma=1.25, r_sigma=2.0):
    c = O.DosConeC()
    c.half_deg, c.max_packing, c.ui_weight, c.initial_step = half, packing, weight, init
    c.covered, c.d_sigma, c.r_sigma = covered, d_sigma, r_sigma
    return c


SPECS = [spec(), spec(half=0.5, packing=0, weight=1.0, covered=166.0), spec(half=30.0, packing=2, covered=250.0),
         spec(half=45.0, packing=2, covered=60.0, d_sigma=2.0, r_sigma=1.0), spec(half=10.0, packing=0, covered=80.0),
         spec(half=60.0, packing=1, covered=40.0, init=1.5)]


@pytest.mark.parametrize("c", SPECS)
def test_section_table_structure(c):
    sigma0 = 1.0
    rows, counts, axes, adj7 = O.dos_sections(c, sigma0)
    interval, level, dint, amp, pos, radius, sigma, ng = rows.T
    n = len(rows)
    div = {1: 1.0, 3: DIV3, 7: DIV7}
    # positions: first at the initial step, then chained by the intervals; the last
    # interval is 0 and the table stops at the first section past the covered distance
    assert pos[0] == c.initial_step
    assert np.allclose(pos[1:], pos[:-1] + interval[:-1], rtol=0, atol=1e-9)
    assert interval[-1] == 0.0
    assert pos[-1] >= c.covered and pos[-2] < c.covered
    # sigma: powers of two of sigma0, never decreasing; the level is their log2
    k = np.log2(sigma / sigma0)
    assert np.array_equal(k, np.round(k)) and np.all(np.diff(sigma) >= 0)
    assert np.array_equal(level, k)
    # each interval is d_sigma * sigma of the section before plus d_sigma * sigma of the next
    assert np.allclose(interval[:-1], c.d_sigma * (sigma[:-1] + sigma[1:]), rtol=0, atol=1e-12)
    # packings only grow, within the limit; counts agree; the radius is the sub-cone's and fits
    assert np.all(np.diff(ng) >= 0) and ng.max() <= [1, 3, 7][c.max_packing]
    assert list(counts) == [int((ng == 1).sum()), int((ng == 3).sum()), int((ng == 7).sum())]
    want_r = pos * np.tan(np.radians(c.half_deg / np.vectorize(div.get)(ng)))
    assert np.allclose(radius, want_r, rtol=1e-12, atol=0)
    assert np.all(radius <= c.r_sigma * sigma * (1 + 1e-12))
    # trapezoid half widths: the first covers a Gaussian's width, then half the previous interval
    assert dint[0] == pytest.approx(sigma[0] * math.sqrt(2 * math.pi) * 0.5, rel=1e-15)
    assert np.allclose(dint[1:], 0.5 * interval[:-1], rtol=1e-15)
    # amplitude = (mass over the diameter * Gaussian integral)^2 / disc area, the mass by
    # the erf closed form; the table integrates with the midpoint rule on segments of at
    # most 0.05, whose error is below 2R h^2 max|f''| / 24 with max|f''| = 1/(sqrt(2 pi) s^3)
    mass = np.array([math.erf(r / (s * math.sqrt(2))) for r, s in zip(radius, sigma)])
    want = (mass * sigma * math.sqrt(2 * math.pi)) ** 2 / (math.pi * radius ** 2)
    h = 2 * radius / np.ceil(2 * radius / 0.05)
    bound = 2 * radius * h ** 2 / 24 / (math.sqrt(2 * math.pi) * sigma ** 3) / mass
    rel = np.abs(amp / want - 1)
    assert np.all(rel <= 2.01 * bound + 1e-12), np.max(rel / bound)
    assert rel.max() > 0 or n < 3  # the table does not use the closed form
    assert n == counts.sum()


@pytest.mark.parametrize("half", [20.0, 30.0, 5.0])
def test_sub_cone_axes(half):
    _, _, axes, adj7 = O.dos_sections(spec(half=half, packing=2), 1.0)
    assert np.allclose(np.linalg.norm(axes, axis=1), 1.0, atol=1e-14)
    t3, t7 = math.radians(half - half / DIV3), math.radians(half - half / DIV7)
    for j in range(3):  # polar t3, azimuths 30, 150, 270 degrees
        phi = math.radians(30 + 120 * j)
        want = [math.sin(t3) * math.cos(phi), math.sin(t3) * math.sin(phi), math.cos(t3)]
        assert np.allclose(axes[j], want, atol=1e-12)
    assert np.allclose(axes[3], [0, 0, 1], atol=0)
    for j in range(1, 7):  # polar t7, azimuths 0, 60, ..., 300 degrees
        phi = math.radians(60 * (j - 1))
        want = [math.sin(t7) * math.cos(phi), math.sin(t7) * math.sin(phi), math.cos(t7)]
        assert np.allclose(axes[3 + j], want, atol=1e-12)
    assert adj7 == pytest.approx(math.cos(t7), abs=1e-14)


def uniform_pyramid(tau, grid, levels=5, n=4):
    return O.OPyr.wrap(grid, [np.full((n, n, n), tau) for _ in range(levels)])


@pytest.mark.parametrize("c", [spec(), spec(half=30.0, packing=2, covered=150.0), spec(half=0.5, packing=0, weight=1.0)])
def test_uniform_medium_closed_form(c):
    """Constant extinction tau inside a box the cone never leaves: every sub-cone sees
    tau * amplitude at each section, so I = exp(-w tau sum_i amp_i (d_i + d_{i+1}))."""
    tau = 0.013
    grid = (600.0, 600.0, 600.0)
    pyr = uniform_pyramid(tau, grid)
    rows = O.dos_sections(c)[0]
    amp, dint = rows[:, 3], rows[:, 2]
    depth = c.ui_weight * tau * (np.sum(amp * dint) + np.sum(amp[:-1] * dint[1:]))
    k = np.array([0.3, -0.5, 0.8]); k /= np.linalg.norm(k)
    u = np.cross(k, [1.0, 0, 0]); u /= np.linalg.norm(u)
    v = np.cross(k, u)
    got = O.dos_cone(pyr, c, [300.0, 300.0, 300.0], k, u, v)
    assert got == pytest.approx(math.exp(-depth), rel=1e-12)
    assert got < 0.99


def test_cone_sees_only_its_half_space():
    """Extinction only at z above the centre: cones opening towards -z see nothing (I = 1
    exactly), towards +z they are occluded, and rotating the frame about +z changes
    nothing for a medium that is invariant in x and y."""
    grid = (400.0, 400.0, 400.0)
    n = 8
    lv = np.zeros((n, n, n))
    lv[n // 2 + 1:, :, :] = 0.05
    pyr = O.OPyr.wrap(grid, [lv] * 5)
    c = spec(half=30.0, packing=2, covered=150.0)
    p = [200.0, 200.0, 200.0]
    ez, ex, ey = np.array([0, 0, 1.0]), np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    assert O.dos_cone(pyr, c, p, -ez, ey, ex) == 1.0
    a = O.dos_cone(pyr, c, p, ez, ey, ex)
    assert a < 0.5
    cs, sn = math.cos(0.7), math.sin(0.7)
    b = O.dos_cone(pyr, c, p, ez, cs * ey + sn * ex, cs * ex - sn * ey)
    assert b == pytest.approx(a, rel=1e-12)


def test_dos_zero_medium_equals_plain_march():
    """With an empty pyramid every cone gives I = 1, so the occlusion-only shading leaves
    the TF colour unchanged and the image is the plain rc1pass march."""
    from cpp_volume_rendering_amd import native as N
    from cpp_volume_rendering_amd import workloads as W
    wl = W.plain_blobs(n=24, width=20, height=16, seed=5)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    pyr = O.OPyr.wrap(wl.grid, [np.zeros((24 >> i,) * 3) for i in range(4)])
    p = N.DosParams()
    p.apply_occlusion, p.apply_shadow, p.ka, p.kd = 1, 0, 0.5, 0.5
    for c in (p.occ, p.sdw):
        c.half_deg, c.max_packing, c.ui_weight, c.initial_step, c.covered_dist, c.d_sigma, c.r_sigma = (
            20.0, 2, 0.35, 3.0, 0.0, 1.25, 2.0)
    prm = O.dos_params(p, wl.grid)
    cam = wl.camera()
    got, _ = O.render_dos(wl.volume, tf, pyr, prm, cam, 20, 16)
    want, _ = O.render_rc1p(wl.volume, tf, cam, 20, 16)
    assert want[..., 3].max() > 0.1
    assert np.max(np.abs(got - want)) <= 1e-12


# ---------------------------------------------------------------------------- ground truth
def minstd_canonical(n):
    """libstdc++'s std::default_random_engine is minstd_rand0 (x <- 16807 x mod 2^31-1,
    seed 1); uniform_real_distribution<float>(0,1) maps a draw x to (x - 1) / (2^31 - 2)
    in float (generate_canonical with one draw, clamped below 1)."""
    x, out = 1, []
    for _ in range(n):
        x = (16807 * x) % 2147483647
        u = np.float32(np.float32(x - 1) / np.float32(2147483646.0))
        out.append(min(u, np.nextafter(np.float32(1), np.float32(0))))
    return out, x


def test_minstd_known_answer():
    # the C++ standard fixes the 10000th output of a default-constructed minstd_rand0
    _, x = minstd_canonical(10000)
    assert x == 1043618065


def test_gt_ray_tables_follow_the_engine_stream():
    g = O.GtC()
    g.n_occ, g.n_sdw, g.occ_aperture_deg, g.sdw_aperture_deg = 5, 3, 90.0, 1.0
    occ, sdw = O.gt_rays(g)
    u, _ = minstd_canonical(2 * (5 + 3))
    pi = np.float32(np.pi)

    def ray(i, aperture):
        th = float(np.float32(u[2 * i]) * np.float32(pi * np.float32(np.float32(aperture) / np.float32(180.0))))
        ph = float(np.float32(u[2 * i + 1]) * np.float32(pi * np.float32(np.float32(360.0) / np.float32(180.0))))
        return [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]

    for i in range(5):
        assert np.allclose(occ[i], ray(i, 90.0), atol=1e-12)
    for i in range(3):
        assert np.allclose(sdw[i], ray(5 + i, 1.0), atol=1e-12)
    ang = np.degrees(np.arccos(np.clip(occ[:, 2], -1, 1)))
    assert np.all(ang <= 90.0 + 1e-9)


def const_volume(n, value):
    return np.full((n, n, n), value, np.uint8)


def test_gt_ray_uniform_closed_form():
    """Constant density: the trapezoid rule is exact, so Vt = exp(-tau L) with L the
    distance of whole steps taken before a step would leave the box (or the march
    length); a dense medium stops at the first step past opacity 0.99."""
    n = 32
    vol = const_volume(n, 128)
    tf = O.OTF([[1, 1, 1, 0], [1, 1, 1, 255]], [[0.0, 0], [0.05, 255]])
    rgbt = tf.tables()[1]
    tau = np.zeros(4)
    O.lib().orc_tex1d(rgbt.ctypes.data, tf.T, 128 / 255.0, tau.ctypes.data)
    tau = tau[3]
    p = np.array([10.2, 16.0, 16.0])
    # along +x: the box face is 21.8 away; march length 100 -> whole steps up to the face
    gap, h = 1.0, 0.5
    E = 32.0 - 10.2
    L = math.floor((E - gap) / h) * h
    got = O.gt_ray(vol, tf, p, [1.0, 0, 0], 100.0, gap, h)
    assert got == pytest.approx(math.exp(-tau * L), rel=1e-12)
    # march length shorter than the face distance: the last step is partial
    got = O.gt_ray(vol, tf, p, [1.0, 0, 0], 7.3, gap, h)
    assert got == pytest.approx(math.exp(-tau * (7.3 - gap)), rel=1e-12)
    # dense medium: stops at the first whole step with 1 - Vt > 0.99
    tf2 = O.OTF([[1, 1, 1, 0], [1, 1, 1, 255]], [[0.0, 0], [0.9, 255]])
    t2 = np.zeros(4)
    O.lib().orc_tex1d(tf2.tables()[1].ctypes.data, tf2.T, 128 / 255.0, t2.ctypes.data)
    m = math.ceil(math.log(100.0) / (t2[3] * h))
    got = O.gt_ray(vol, tf2, p, [1.0, 0, 0], 100.0, gap, h)
    assert m * h < E - gap
    assert got == pytest.approx(math.exp(-t2[3] * h * m), rel=1e-12)


def test_gt_zero_aperture_cone_is_its_axis():
    """With a zero aperture every sampled ray is the cone axis, so the cosine-weighted
    mean over any number of rays equals one ray."""
    from cpp_volume_rendering_amd import native as N
    from cpp_volume_rendering_amd import workloads as W
    wl = W.cone_blobs(n=24, width=16, height=12, seed=9)
    tf = O.OTF(wl.rgb_iso, wl.alpha_iso)
    p = N.GtParams()
    p.apply_occlusion, p.apply_shadow, p.occ_rays, p.sdw_rays = 1, 1, 1, 1
    p.occ_aperture_deg, p.sdw_aperture_deg, p.shadow_type = 0.0, 0.0, 0
    p.initial_gap, p.ray_step, p.ka, p.kd = 1.0, 0.5, 0.5, 0.5
    light = wl.light()
    one, _ = O.render_gt(wl.volume, tf, O.gt_params(p, wl.grid, light), wl.camera(), 16, 12)
    p.occ_rays, p.sdw_rays = 6, 4
    many, _ = O.render_gt(wl.volume, tf, O.gt_params(p, wl.grid, light), wl.camera(), 16, 12)
    assert one[..., 3].max() > 0.1
    assert np.max(np.abs(one - many)) <= 1e-12
