"""The Python binding checks output buffers and volume types before any address reaches
the C ABI: a wrong dtype, shape or layout, or a host torch tensor where device memory is
expected, raises CvrError instead of letting a kernel or a copy write out of bounds (CPU
only: no call reaches the GPU)."""
import numpy as np
import pytest
import torch

from cpp_volume_rendering_amd import native as N


def test_numpy_output_checked():
    good = np.zeros((4, 5, 4), np.float32)
    ptr, dev = N._out_arg(good, (4, 5, 4))
    assert ptr == good.ctypes.data and dev is False
    for bad in (np.zeros((4, 5, 4), np.float64), np.zeros((4, 4, 4), np.float32), np.zeros((80,), np.float32),
                np.zeros((5, 4, 4), np.float32).transpose(1, 0, 2)):
        with pytest.raises(N.CvrError) as e:
            N._out_arg(bad, (4, 5, 4))
        assert e.value.code == N.ERR_INVALID


def test_torch_output_must_be_device_float32_contiguous():
    with pytest.raises(N.CvrError, match="GPU"):
        N._out_arg(torch.zeros((4, 5, 4)), (4, 5, 4))  # a host tensor would get a host address
    with pytest.raises(N.CvrError):
        N._out_arg(torch.zeros((4, 5, 4), dtype=torch.float16), (4, 5, 4))
    with pytest.raises(N.CvrError):
        N._out_arg(torch.zeros((5, 4, 4)).transpose(0, 1), (4, 5, 4))
    with pytest.raises(N.CvrError):
        N._out_arg(torch.zeros((2, 5, 4)), (4, 5, 4))
    with pytest.raises(N.CvrError):
        N._out_arg([0.0] * 80, (4, 5, 4))


def test_volume_dtypes():
    assert N._volume_dtype("uint8") == N.U8 and N._volume_dtype("torch.uint8") == N.U8
    assert N._volume_dtype("uint16") == N.U16 and N._volume_dtype("torch.uint16") == N.U16
    for name in ("float32", "int16", "int32", "torch.int16", "torch.float32"):
        with pytest.raises(N.CvrError):
            N._volume_dtype(name)
