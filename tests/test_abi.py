"""The C-ABI library loads, exports every symbol the headers declare, and its
argument checks fail cleanly without a GPU (CPU only)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from cpp_volume_rendering_amd import native as N

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(cvr_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_all_declared_symbols():
    L = N.lib()
    names = declared_symbols()
    assert names, "no declarations found"
    for n in sorted(names):
        assert hasattr(L, n), f"libcvr.so lacks {n}"
    assert set(N.EXPORTS) <= names


def test_default_params_and_names():
    L = N.lib()
    p = N.EbsParams()
    N.check(L.cvr_default_params(N.EBS, C.byref(p), C.sizeof(p)))
    assert (p.ao_shells, p.ao_radius, p.sdw_angle_deg, p.sdw_interval, p.sdw_initial) == (15, 1.0, 1.0, 2.0, 2.0)
    assert (p.ka, p.kd, p.et_threshold) == (0.5, 0.5, pytest.approx(0.99))
    assert L.cvr_default_params(N.EBS, C.byref(p), C.sizeof(p) - 4) == N.ERR_INVALID
    assert L.cvr_technique_name(N.RC1P) == b"1-Pass - Ray Casting"
    assert L.cvr_technique_name(99) is None


def test_create_argument_errors():
    L = N.lib()
    assert L.cvr_create(N.RC1P, 0, None) == N.ERR_INVALID
    h = C.c_void_p()
    assert L.cvr_create(99, 0, C.byref(h)) == N.ERR_UNSUPPORTED
    # null handles never crash
    assert L.cvr_set_volume(None, 4, 4, 4, 0, 1, 1, 1, None, 0) == N.ERR_INVALID
    assert L.cvr_render(None, None, None, 4, 4, None, 0, None) == N.ERR_INVALID
    L.cvr_destroy(None)


def test_synthetic_generators_deterministic():
    a = N.synth_blobs((20, 18, 16), N.U8, n_blobs=5, seed=9)
    b = N.synth_blobs((20, 18, 16), N.U8, n_blobs=5, seed=9)
    c = N.synth_blobs((20, 18, 16), N.U8, n_blobs=5, seed=10)
    assert a.shape == (16, 18, 20) and a.max() == 255
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    n = N.synth_noise((24, 24, 24), N.U16, octaves=3, seed=2, bands=[0.2, 0.5, 0.8])
    vals, counts = np.unique(n, return_counts=True)
    assert list(vals) == [13107, 32768, 52428]
    assert counts.min() > 0.3 * n.size
