"""ctypes access to the CPU oracle (oracle/_build/liboracle.so) — test infrastructure.

Builds the oracle with its Makefile when the shared object is missing (g++ only,
no GPU).  Only tests, smoke() and bench.py's CPU-baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
SO = ORACLE_DIR / "_build" / "liboracle.so"

_lib = None


class EbsC(C.Structure):
    _fields_ = [("apply_occlusion", C.c_int), ("apply_shadow", C.c_int), ("ao_shells", C.c_int),
                ("shadow_type", C.c_int), ("ao_radius", C.c_float), ("sdw_angle_deg", C.c_float),
                ("sdw_interval", C.c_float), ("sdw_initial", C.c_float), ("sdw_weight", C.c_float),
                ("sdw_max_dist", C.c_float), ("light_pos", C.c_float * 3), ("light_fwd", C.c_float * 3),
                ("ka", C.c_float), ("kd", C.c_float)]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not SO.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
    L = C.CDLL(str(SO))
    vp, i, d, f = C.c_void_p, C.c_int, C.c_double, C.c_float
    sig = {
        "orc_max_threads": (i, []),
        "orc_tf_new": (vp, [i, i, i, vp, i, vp]),
        "orc_tf_free": (None, [vp]),
        "orc_tf_size": (i, [vp]),
        "orc_tf_tables": (None, [vp, vp, vp, vp]),
        "orc_tf_get": (None, [vp, d, d, vp]),
        "orc_tf_get_ext_n": (f, [vp, d]),
        "orc_tex1d": (None, [vp, i, d, vp]),
        "orc_tex3d_volume": (d, [i, i, i, i, vp, d, d, d]),
        "orc_tex3d_grid": (d, [vp, i, i, i, d, d, d]),
        "orc_rays": (None, [vp, i, i, d, d, d, vp]),
        "orc_render_rc1p_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, i, i, d, d, vp, vp, i, i]),
        "orc_render_rc1p_f32": (None, [i, i, i, i, vp, d, d, d, vp, vp, i, i, f, f, vp, i]),
        "orc_sat_f64": (None, [i, i, i, i, vp, vp, vp, i]),
        "orc_sat_input_f64": (None, [i, i, i, i, vp, vp, vp]),
        "orc_render_ebs_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, vp, i, i, d, d, vp, vp, i, i]),
        "orc_render_ebs_f32": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, vp, i, i, f, f, vp, i]),
        "orc_ebs_point": (None, [i, i, i, d, d, d, vp, vp, d, d, d, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def cam_array(cam) -> np.ndarray:
    """native.Camera -> float32[11] (eye, center, up, fovy, aspect)."""
    return np.array(list(cam.eye) + list(cam.center) + list(cam.up) + [cam.fovy_deg, cam.aspect], dtype=np.float32)


class OTF:
    """Oracle transfer function handle."""

    def __init__(self, rgb_iso, alpha_iso, max_density=255, extinction_input=False):
        self.rgb = np.ascontiguousarray(rgb_iso, dtype=np.float64).reshape(-1, 4)
        self.a = np.ascontiguousarray(alpha_iso, dtype=np.float64).reshape(-1, 2)
        self.h = lib().orc_tf_new(max_density, int(extinction_input), len(self.rgb), self.rgb.ctypes.data,
                                  len(self.a), self.a.ctypes.data)
        if not self.h:
            raise ValueError("invalid transfer function control points")
        self.T = lib().orc_tf_size(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tf_free(self.h)

    def tables(self):
        tab = np.zeros((self.T, 4), np.float64)
        rgbt = np.zeros((self.T, 4), np.float32)
        rgba = np.zeros((self.T, 4), np.float32)
        lib().orc_tf_tables(self.h, tab.ctypes.data, rgbt.ctypes.data, rgba.ctypes.data)
        return tab, rgbt, rgba

    def get(self, value, maxv=-1.0):
        out = np.zeros(4, np.float32)
        lib().orc_tf_get(self.h, value, maxv, out.ctypes.data)
        return out

    def get_ext_n(self, v):
        return lib().orc_tf_get_ext_n(self.h, v)


def _vol_args(vol, scale):
    vol = np.ascontiguousarray(vol)
    nz, ny, nx = vol.shape
    dt = 0 if vol.dtype == np.uint8 else 1
    return vol, (nx, ny, nz, dt, vol.ctypes.data, float(scale[0]), float(scale[1]), float(scale[2]))


def render_rc1p(vol, tf: OTF, cam, W, H, step=0.5, et=0.99, scale=(1, 1, 1), fp32=False, threads=0, stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    if fp32:
        out = np.zeros((H, W, 4), np.float32)
        lib().orc_render_rc1p_f32(*va, tf.h, c.ctypes.data, W, H, step, et, out.ctypes.data, threads)
        return out
    out = np.zeros((H, W, 4), np.float64)
    margin = np.zeros((H, W, 2), np.float64)
    lib().orc_render_rc1p_f64(*va, tf.h, c.ctypes.data, W, H, step, et, out.ctypes.data, margin.ctypes.data, threads,
                              stride)
    return out, margin


def sat(vol, tf: OTF, threads=0):
    vol = np.ascontiguousarray(vol)
    nz, ny, nx = vol.shape
    out = np.zeros((nz + 2, ny + 2, nx + 2), np.float64)
    lib().orc_sat_f64(nx, ny, nz, 0 if vol.dtype == np.uint8 else 1, vol.ctypes.data, tf.h, out.ctypes.data, threads)
    return out


def sat_input(vol, tf: OTF):
    vol = np.ascontiguousarray(vol)
    nz, ny, nx = vol.shape
    out = np.zeros((nz + 2, ny + 2, nx + 2), np.float64)
    lib().orc_sat_input_f64(nx, ny, nz, 0 if vol.dtype == np.uint8 else 1, vol.ctypes.data, tf.h, out.ctypes.data)
    return out


def ebs_params(p=None, light=None, **kw) -> EbsC:
    """From a native.EbsParams (+ native.Light) or keywords, with reference defaults."""
    e = EbsC()
    e.apply_occlusion, e.apply_shadow, e.ao_shells, e.shadow_type = 1, 1, 15, 1
    e.ao_radius, e.sdw_angle_deg, e.sdw_interval, e.sdw_initial, e.sdw_weight = 1.0, 1.0, 2.0, 2.0, 1.0
    e.sdw_max_dist, e.ka, e.kd = 0.0, 0.5, 0.5
    if p is not None:
        for k in ("apply_occlusion", "apply_shadow", "ao_shells", "shadow_type", "ao_radius", "sdw_angle_deg",
                  "sdw_interval", "sdw_initial", "sdw_weight", "sdw_max_dist", "ka", "kd"):
            setattr(e, k, getattr(p, k))
    if light is not None:
        e.light_pos[:] = list(light.position)
        e.light_fwd[:] = list(light.forward)
    for k, v in kw.items():
        if k in ("light_pos", "light_fwd"):
            getattr(e, k)[:] = list(v)
        else:
            setattr(e, k, v)
    return e


def render_ebs(vol, tf: OTF, sat_arr, prm: EbsC, cam, W, H, step=0.5, et=0.99, scale=(1, 1, 1), fp32=False, threads=0,
               stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    if fp32:
        s32 = np.ascontiguousarray(sat_arr, dtype=np.float32)
        out = np.zeros((H, W, 4), np.float32)
        lib().orc_render_ebs_f32(*va, tf.h, s32.ctypes.data, C.byref(prm), c.ctypes.data, W, H, step, et,
                                 out.ctypes.data, threads)
        return out
    s64 = np.ascontiguousarray(sat_arr, dtype=np.float64)
    out = np.zeros((H, W, 4), np.float64)
    margin = np.zeros((H, W, 2), np.float64)
    lib().orc_render_ebs_f64(*va, tf.h, s64.ctypes.data, C.byref(prm), c.ctypes.data, W, H, step, et,
                             out.ctypes.data, margin.ctypes.data, threads, stride)
    return out, margin


def ebs_point(shape, scale, sat_arr, prm: EbsC, pos):
    nx, ny, nz = shape
    s64 = np.ascontiguousarray(sat_arr, dtype=np.float64)
    ao, sd = C.c_double(), C.c_double()
    lib().orc_ebs_point(nx, ny, nz, *map(float, scale), s64.ctypes.data, C.byref(prm), *map(float, pos),
                        C.byref(ao), C.byref(sd))
    return ao.value, sd.value


# Straddle rule shared by the parity tests: a pixel whose FP64 march passed a
# threshold decision closer than these margins may legitimately take the other
# branch in fp32.  Alpha/transmittance: 2e-5 (fp32 accumulated alpha is good to
# ~1e-6).  Geometry: 5e-4 world units (fp32 sample positions carry ~1.5e-4 of
# rounding at eye distances of ~900 voxels).
MARGIN_ALPHA = 2e-5
MARGIN_GEOM = 5e-4


def compare(img_gpu, img_ref, margin=None, tol=2e-3, stride=1):
    """Parity summary: max-abs per channel over non-straddle pixels, RMSE, counts.
    With stride > 1 only the pixels the oracle rendered (x, y multiples of stride) count."""
    g = np.asarray(img_gpu, np.float64)[::stride, ::stride]
    r = np.asarray(img_ref, np.float64)[::stride, ::stride]
    if margin is not None:
        margin = margin[::stride, ::stride]
    diff = np.abs(g - r)
    strad = np.zeros(diff.shape[:2], bool)
    if margin is not None:
        strad = (margin[..., 0] < MARGIN_ALPHA) | (margin[..., 1] < MARGIN_GEOM)
    ok = ~strad
    per_ch = diff[ok].max(axis=0) if ok.any() else np.zeros(4)
    return {
        "max_abs": float(per_ch.max()),
        "max_abs_rgba": [float(x) for x in per_ch],
        "max_abs_all": float(diff.max()),
        "rmse": float(np.sqrt((diff ** 2).mean())),
        "n_over_tol": int((diff.max(axis=2) > tol).sum()),
        "n_over_tol_nonstraddle": int(((diff.max(axis=2) > tol) & ok).sum()),
        "n_straddle": int(strad.sum()),
        "n_pixels": int(diff.shape[0] * diff.shape[1]),
    }
