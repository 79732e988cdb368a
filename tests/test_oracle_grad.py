"""Pins the oracle's volume gradients and Blinn-Phong shading (CPU only): closed forms
on linear ramps, a numpy restatement with zero padding at the borders, a numpy march
of single pixels with ShadeBlinnPhong, and the identity between the lit Phong branch
(unit shadow, no occlusion) and the plain one."""
import math

import numpy as np
import pytest

import oracle_lib as O
from cpp_volume_rendering_amd import native as N
from cpp_volume_rendering_amd import workloads as W


def ramp(shape=(9, 8, 7), coef=(300, 700, 1100), base=2000):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (base + coef[0] * x + coef[1] * y + coef[2] * z).astype(np.uint16)


def numpy_sobel(v):
    s = np.pad(v.astype(np.float64) / 65535.0, 1)  # 0 outside the volume
    nz, ny, nx = v.shape
    g = np.zeros(v.shape + (3,))
    for v1 in (-1, 0, 1):
        for v2 in (-1, 0, 1):
            w = 4.0 / 2 ** (abs(v1) + abs(v2))
            sh = lambda dz, dy, dx: s[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
            g[..., 2] += sh(-1, v2, v1) * w - sh(1, v2, v1) * w
            g[..., 1] += sh(v2, -1, v1) * w - sh(v2, 1, v1) * w
            g[..., 0] += sh(v1, v2, -1) * w - sh(v1, v2, 1) * w
    return g


def test_sobel_ramp_closed_form_and_borders():
    v = ramp()
    g = O.gradient(v, 0)
    # interior: each axis sums 16 weights times (s(-1) - s(+1)) = -2 * slope
    want = -32.0 * np.array([300, 700, 1100]) / 65535.0
    assert np.allclose(g[1:-1, 1:-1, 1:-1], want, rtol=0, atol=1e-12)
    # everywhere, borders included (samples outside are 0)
    assert np.allclose(g, numpy_sobel(v), rtol=0, atol=1e-12)


def test_finite_differences_normalised_and_zero():
    v = ramp()
    g = O.gradient(v, 1)
    d = np.array([300, 700, 1100], float)
    assert np.allclose(g[1:-1, 1:-1, 1:-1], d / np.linalg.norm(d), atol=1e-12)
    c = np.full((5, 6, 7), 40, np.uint8)
    gc = O.gradient(c, 1)
    assert np.all(gc[1:-1, 1:-1, 1:-1] == 0.0)  # vanishing difference -> the zero vector
    assert np.allclose(np.linalg.norm(gc[0, 2, 2]), 1.0)  # the border sees the 0 outside


def numpy_phong_pixel(vol, tf, grad, ph, cam, W_, H_, px, py, step=0.5, et=0.99):
    nz, ny, nx = vol.shape
    G = np.array([nx, ny, nz], float)
    eye, ce, up = (np.array(v, float) for v in (cam.eye, cam.center, cam.up))
    f = (ce - eye) / np.linalg.norm(ce - eye)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    t = float(np.float32(math.tan(math.radians(cam.fovy_deg) / 2)))
    asp = cam.aspect if cam.aspect > 0 else float(np.float32(W_ / H_))
    d = s * ((px + 0.5) / W_ * 2 - 1) * t * asp + u * ((py + 0.5) / H_ * 2 - 1) * t + f
    d /= np.linalg.norm(d)
    with np.errstate(divide="ignore"):
        t0, t1 = (-G / 2 - eye) / d, (G / 2 - eye) / d
    tn, tf_ = max(np.minimum(t0, t1).max(), 0.0), np.maximum(t0, t1).min()
    dst = np.zeros(4)
    if not tf_ > tn:
        return dst
    rgbt = np.ascontiguousarray(tf.tables()[1])

    def trilinear(p):
        c = p / G * np.array([nx, ny, nz]) - 0.5
        i0 = np.floor(c).astype(int)
        fr = c - i0
        lo, hi = np.clip(i0, 0, [nx - 1, ny - 1, nz - 1]), np.clip(i0 + 1, 0, [nx - 1, ny - 1, nz - 1])
        acc = np.zeros(3)
        for k in range(8):
            sel = [(k >> a) & 1 for a in range(3)]
            idx = [hi[a] if sel[a] else lo[a] for a in range(3)]
            acc += np.prod([fr[a] if sel[a] else 1 - fr[a] for a in range(3)]) * grad[idx[2], idx[1], idx[0]]
        return acc

    lpos = np.array(ph.lpos[:])
    D, tp, sp_ = abs(tf_ - tn), eye + d * tn + G / 2, 0.0
    while sp_ < D:
        h = min(step, D - sp_)
        p = tp + d * (sp_ + h / 2)
        dens = O.lib().orc_tex3d_volume(nx, ny, nz, 0 if vol.dtype == np.uint8 else 1, vol.ctypes.data, *map(float, p / G))
        src = np.zeros(4)
        O.lib().orc_tex1d(rgbt.ctypes.data, tf.T, float(dens), src.ctypes.data)
        if src[3] > 0:
            n = trilinear(p)
            if np.any(n != 0):
                n /= np.linalg.norm(n)
                wp = p - G / 2
                L = (lpos - wp) / np.linalg.norm(lpos - wp)
                E = (eye - wp) / np.linalg.norm(eye - wp)
                Hh = (E + L) / np.linalg.norm(E + L)
                dd, ds = max(0.0, n @ L), max(0.0, Hh @ n)
                src[:3] = src[:3] * (ph.ka + ph.kd * dd) + np.array(ph.spec[:]) * ph.ks * ds ** ph.shininess
            a = 1 - math.exp(-src[3] * h)
            dst[:3] += (1 - dst[3]) * src[:3] * a
            dst[3] += (1 - dst[3]) * a
            if dst[3] > et:
                break
        sp_ += h
    return dst


@pytest.mark.parametrize("kind", [0, 1])
def test_plain_phong_matches_a_numpy_march(kind):
    n = 12
    vol = np.ascontiguousarray(N.synth_blobs((n, n, n), N.U8, n_blobs=5, seed=31))
    rgb, a = W.seeded_tf(31, 4)
    tf = O.OTF(rgb, a)
    cam = W.orbit_camera((float(n),) * 3, 16, 12, 40.0, 20.0, dist_diag=0.9)
    ph = O.phong(kind, lpos=(20.0, 30.0, 25.0), ks=0.8, shininess=12.0)
    img, _ = O.render_rc1p_phong(vol, tf, ph, cam, 16, 12)
    grad = O.gradient(vol, kind)
    lit = 0
    for py in range(0, 12, 2):
        for px in range(0, 16, 2):
            ref = numpy_phong_pixel(vol, tf, grad, ph, cam, 16, 12, px, py)
            assert np.abs(img[py, px] - ref).max() < 1e-9, (px, py)
            lit += ref[3] > 0.01
    assert lit > 8
    plain, _ = O.render_rc1p(vol, tf, cam, 16, 12)
    assert np.abs(img - plain).max() > 1e-3  # the shading does something


def test_lit_phong_reduces_to_the_plain_phong():
    """The lit Phong branch (shared by the EBS, DOS and light-cache marchers) with unit
    shadow, no occlusion and kd = 1 is ShadeBlinnPhong with ka = 0, kd = 1."""
    n = 16
    vol = N.synth_blobs((n, n, n), N.U8, n_blobs=5, seed=32)
    rgb, a = W.seeded_tf(32, 4)
    tf = O.OTF(rgb, a)
    cam = W.orbit_camera((float(n),) * 3, 12, 12, 30.0, 20.0, dist_diag=1.0)
    for kind in (0, 1):
        ph = O.phong(kind, lpos=(10.0, 40.0, 5.0))
        lc, _ = O.render_lc_phong(vol, tf, np.ones((4, 4, 4, 2)), 0, 1, 0.5, 1.0, ph, cam, 12, 12)
        rp, _ = O.render_rc1p_phong(vol, tf, O.phong(kind, lpos=(10.0, 40.0, 5.0), ka=0.0, kd=1.0), cam, 12, 12)
        assert np.abs(lc - rp).max() < 1e-12
        plain, _ = O.render_rc1p(vol, tf, cam, 12, 12)
        assert np.abs(rp - plain).max() > 1e-3
