"""Pins the oracle's multi-scale output resampling (CPU only): kernel closed forms and
partitions of unity, the digital filter inverting the kernel's sampled convolution,
box downscaling = block means, multisampling at half resolution = block means, and the
interpolating kernels reproducing their samples."""
import numpy as np
import pytest

import oracle_lib as O

@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_kernels_partition_of_unity(k):
    """A constant frame upscaled by 3: away from the borders every output pixel is
    c * sum_i k(xi - i), which is c for the hat, Catmull-Rom and Mitchell-Netravali; the
    cardinal kernels sum to 6 resp. 5.25 and their digital filter divides a constant
    by exactly that."""
    c = 0.37
    f = np.full((24, 24, 4), c)
    u = O.resample(f, 72, 72, 2, k)
    assert np.allclose(u[24:-24, 24:-24], c, atol=1e-12 if k < 4 else 2e-7)


def test_mitchell_netravali_values():
    # B = C = 1/3: k(0) = 8/9, k(1) = 1/18, k(2) = 0; seen through an impulse upscaled 1:1
    f = np.zeros((9, 9, 4))
    f[4, 4] = 1.0
    u = O.resample(f, 9, 9, 2, 3)
    assert u[4, 4, 0] == pytest.approx((8 / 9) ** 2, abs=1e-12)
    assert u[4, 5, 0] == pytest.approx((8 / 9) * (1 / 18), abs=1e-12)
    assert u[4, 6, 0] == pytest.approx(0.0, abs=1e-12)


def test_box_downscale_and_multisample_are_block_means():
    f = np.random.default_rng(0).uniform(0, 1, (16, 20, 4))
    mean = f.reshape(8, 2, 10, 2, 4).mean(axis=(1, 3))
    assert np.abs(O.resample(f, 10, 8, 1, 0) - mean).max() < 1e-15
    assert np.abs(O.resample(f, 10, 8, 0, 0) - mean).max() < 1e-15


def test_hat_upscale_of_a_ramp_is_linear():
    x = np.arange(10)
    f = np.zeros((8, 10, 4))
    f[..., 0] = 0.1 * x[None, :] + 0.05 * np.arange(8)[:, None]
    u = O.resample(f, 40, 32, 2, 1)
    xi = (np.arange(40) + 0.5) / 40 * 10 - 0.5
    yi = (np.arange(32) + 0.5) / 32 * 8 - 0.5
    want = 0.1 * xi[None, :] + 0.05 * yi[:, None]
    assert np.abs(u[4:-4, 4:-4, 0] - want[4:-4, 4:-4]).max() < 1e-12


@pytest.mark.parametrize("k", [2, 4, 5])
def test_interpolating_kernels_reproduce_their_samples(k):
    f = np.random.default_rng(1).uniform(0, 1, (24, 30, 4))
    u = O.resample(f, 30, 24, 2, k)  # 1:1 upscale evaluates at the sample points
    tol = 1e-12 if k == 2 else 3e-7  # the digital filter's factors are float, as the reference's
    assert np.abs(u[3:-3, 3:-3] - f[3:-3, 3:-3]).max() < tol


@pytest.mark.parametrize("k,a", [(4, 4.0), (5, 3.25)])
def test_digital_filter_inverts_the_sampled_kernel(k, a):
    """Downscaling 1:1 with a cardinal kernel: the kernel's (1, a, 1) taps (plus the
    digital filter) give back the frame in the interior; n <= m uses the reflected
    sequence."""
    f = np.random.default_rng(2).uniform(0, 1, (40, 46, 4))
    d = O.resample(f, 46, 40, 1, k)
    # the zero-padded border and the filter's mirrored boundary differ; their effect
    # decays like L_inf^distance (0.27 resp. 0.34 per pixel)
    assert np.abs(d[16:-16, 16:-16] - f[16:-16, 16:-16]).max() < 1e-6
    small = np.random.default_rng(3).uniform(0, 1, (5, 6, 4))
    ds = O.resample(small, 6, 5, 1, k)
    assert np.all(np.isfinite(ds))
