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
        "orc_render_ebs_f32": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, vp, i, i, f, f, vp, i, i]),
        "orc_ebs_point": (None, [i, i, i, d, d, d, vp, vp, d, d, d, vp, vp]),
        "orc_pyr_new": (vp, [i, i, i, i, vp, d, d, d, vp, i, i, i, d, i, i]),
        "orc_pyr_wrap": (vp, [d, d, d, i, vp, vp]),
        "orc_pyr_free": (None, [vp]),
        "orc_pyr_levels": (i, [vp]),
        "orc_pyr_dims": (None, [vp, i, vp]),
        "orc_pyr_copy": (None, [vp, i, vp, vp]),
        "orc_pyr_level0_at": (d, [i, i, i, i, vp, d, d, d, vp, i, i, i, d, i, i, i]),
        "orc_pyr_next_at": (d, [vp, i, i, i, i, i, i, d, d, d, d, i, i, i]),
        "orc_gauss_ext": (d, [vp, d, d, d, d]),
        "orc_cone_schedule": (None, [i, d, d, d, d, d, i, vp]),
        "orc_render_cone_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, vp, i, i, d, d, vp, vp, i, i]),
        "orc_render_cone_f32": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, vp, i, i, f, f, vp, i, i]),
        "orc_cone_points": (None, [vp, vp, i, vp, vp, vp, vp, vp, i]),
        "orc_dos_sections": (i, [vp, d, vp, i, vp, vp, vp]),
        "orc_dos_cone": (d, [vp, vp, d, vp, vp, vp, vp]),
        "orc_render_dos_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, d, vp, i, i, d, d, vp, vp, i, i]),
        "orc_render_dos_f32": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, d, vp, i, i, f, f, vp, i, i]),
        "orc_gt_rays": (None, [vp, vp, vp]),
        "orc_gt_ray": (d, [i, i, i, i, vp, d, d, d, vp, vp, vp, d, d, d]),
        "orc_render_gt_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, i, i, d, d, vp, vp, i, i]),
        "orc_render_gt_f32": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, i, i, f, f, vp, i, i]),
        "orc_lc_ebs": (None, [i, i, i, d, d, d, vp, vp, i, i, i, vp, i]),
        "orc_lc_dos": (i, [i, i, i, d, d, d, vp, vp, d, vp, i, i, i, i, i, vp, i]),
        "orc_render_lc_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, i, i, i, i, i, d, d, vp, i, i, d, d, vp, vp, i,
                                     i]),
        "orc_render_lc_f32": (None, [i, i, i, i, vp, d, d, d, vp, vp, i, i, i, i, i, d, d, vp, i, i, f, f, vp, i, i]),
        "orc_gradient": (None, [i, i, i, i, vp, i, vp, i]),
        "orc_render_rc1p_phong_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, i, i, d, d, vp, vp, i, i]),
        "orc_render_ebs_phong_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, vp, vp, i, i, d, d, vp, vp, i, i]),
        "orc_render_dos_phong_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, d, vp, vp, i, i, d, d, vp, vp, i, i]),
        "orc_render_lc_phong_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, i, i, i, i, i, d, d, vp, vp, i, i, d, d, vp,
                                           vp, i, i]),
        "orc_resample": (None, [vp, i, i, vp, i, i, i, i]),
        "orc_mlp_nparams": (i, []),
        "orc_mlp_forward": (None, [vp, i, vp, vp]),
        "orc_mlp_loss_grad": (d, [vp, vp, vp, i, vp]),
        "orc_mlp_forward_bf16": (None, [vp, i, vp, vp]),
        "orc_render_nn_f64": (None, [i, i, i, i, vp, d, d, d, vp, vp, vp, vp, vp, vp, i, vp, i, i, d, d, vp, vp, i, i,
                                     vp]),
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
                                 out.ctypes.data, threads, stride)
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


# ---------------------------------------------------------------------------- cones
class ConeC(C.Structure):
    _fields_ = [("n_steps", C.c_int), ("n_occ", C.c_int), ("apply_occ", C.c_int), ("apply_sdw", C.c_int),
                ("occ_half_deg", C.c_double), ("ring_deg", C.c_double), ("sdw_half_deg", C.c_double),
                ("t0", C.c_double), ("occ_max", C.c_double), ("sdw_max", C.c_double), ("r_sigma", C.c_double),
                ("sigma0", C.c_double), ("ka", C.c_double), ("kd", C.c_double), ("light_fwd", C.c_double * 3)]


def cone_params(grid=None, light=None, **kw) -> ConeC:
    """Defaults of the cone-traced configuration: 8 occlusion cones of 30 degrees (ring at
    60), one 10-degree shadow cone, 64 steps from t0 = 1 to 0.5 / 0.75 of the diagonal."""
    c = ConeC()
    c.n_steps, c.n_occ, c.apply_occ, c.apply_sdw = 64, 8, 1, 1
    c.occ_half_deg, c.ring_deg, c.sdw_half_deg, c.t0 = 30.0, 60.0, 10.0, 1.0
    c.r_sigma, c.sigma0, c.ka, c.kd = 2.0, 1.0, 0.5, 0.5
    if grid is not None:
        diag = float(np.float32(np.sqrt(np.sum(np.square(np.asarray(grid, np.float32))))))
        c.occ_max, c.sdw_max = 0.5 * diag, 0.75 * diag
    if light is not None:
        c.light_fwd[:] = list(light.forward)
    for k, v in kw.items():
        if k == "light_fwd":
            c.light_fwd[:] = list(v)
        else:
            setattr(c, k, v)
    return c


class OPyr:
    """Oracle extinction pyramid: built from a volume, or wrapping given levels."""

    def __init__(self, h):
        self.h = h
        self.levels = lib().orc_pyr_levels(h)

    @classmethod
    def build(cls, vol, tf: OTF, base=None, sigma0=1.0, nlev=0, scale=(1, 1, 1), threads=0):
        vol, va = _vol_args(vol, scale)
        nz, ny, nx = vol.shape
        bx, by, bz = base if base is not None else (nx, ny, nz)
        return cls(lib().orc_pyr_new(*va, tf.h, bx, by, bz, sigma0, nlev, threads))

    @classmethod
    def wrap(cls, grid, levels):
        """levels: list of (nz, ny, nx) float64 extinction arrays."""
        dims = np.array([[a.shape[2], a.shape[1], a.shape[0]] for a in levels], np.int32).ravel()
        cat = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64).ravel() for a in levels]))
        return cls(lib().orc_pyr_wrap(*map(float, grid), len(levels), dims.ctypes.data, cat.ctypes.data))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pyr_free(self.h)

    def dims(self, l):
        d = np.zeros(3, np.int32)
        lib().orc_pyr_dims(self.h, l, d.ctypes.data)
        return tuple(int(x) for x in d)

    def level(self, l, opacity=False):
        nx, ny, nz = self.dims(l)
        tau = np.zeros((nz, ny, nx), np.float64)
        op = np.zeros((nz, ny, nx), np.float64)
        lib().orc_pyr_copy(self.h, l, tau.ctypes.data, op.ctypes.data)
        return op if opacity else tau

    def ext(self, pos, lod):
        return lib().orc_gauss_ext(self.h, *map(float, pos), float(lod))


def render_cone(vol, tf: OTF, pyr: OPyr, prm: ConeC, cam, W, H, step=0.5, et=0.99, scale=(1, 1, 1), fp32=False,
                threads=0, stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    if fp32:
        out = np.zeros((H, W, 4), np.float32)
        lib().orc_render_cone_f32(*va, tf.h, pyr.h, C.byref(prm), c.ctypes.data, W, H, step, et, out.ctypes.data,
                                  threads, stride)
        return out
    out = np.zeros((H, W, 4), np.float64)
    margin = np.zeros((H, W, 2), np.float64)
    lib().orc_render_cone_f64(*va, tf.h, pyr.h, C.byref(prm), c.ctypes.data, W, H, step, et, out.ctypes.data,
                              margin.ctypes.data, threads, stride)
    return out, margin


def cone_points(pyr: OPyr, prm: ConeC, pos, dirs, threads=0):
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    n = len(pos)
    occ, sdw, feat = np.zeros(n), np.zeros(n), np.zeros((n, 16))
    lib().orc_cone_points(pyr.h, C.byref(prm), n, pos.ctypes.data, dirs.ctypes.data, occ.ctypes.data,
                          sdw.ctypes.data, feat.ctypes.data, threads)
    return occ, sdw, feat


def mlp_forward(prm, X):
    prm = np.ascontiguousarray(prm, np.float64)
    X = np.ascontiguousarray(X, np.float64).reshape(-1, 16)
    Y = np.zeros((len(X), 16))
    lib().orc_mlp_forward(prm.ctypes.data, len(X), X.ctypes.data, Y.ctypes.data)
    return Y


def mlp_loss_grad(prm, X, T):
    prm = np.ascontiguousarray(prm, np.float64)
    X = np.ascontiguousarray(X, np.float64).reshape(-1, 16)
    T = np.ascontiguousarray(T, np.float64).reshape(-1, 2)
    g = np.zeros_like(prm)
    loss = lib().orc_mlp_loss_grad(prm.ctypes.data, X.ctypes.data, T.ctypes.data, len(X), g.ctypes.data)
    return loss, g


def mlp_forward_bf16(prm, X):
    """FP64 forward with a bf16 network's rounding points (inputs, weights and hidden
    activations rounded to bf16, sums in FP64)."""
    prm = np.ascontiguousarray(prm, np.float64)
    X = np.ascontiguousarray(X, np.float64).reshape(-1, 16)
    Y = np.zeros((len(X), 16))
    lib().orc_mlp_forward_bf16(prm.ctypes.data, len(X), X.ctypes.data, Y.ctypes.data)
    return Y


def render_nn(vol, tf: OTF, pyr: OPyr, prm: ConeC, weights, cam, W, H, lo=None, hi=None, bf16=True, step=0.5,
              et=0.99, scale=(1, 1, 1), threads=0, stride=1):
    """The surrogate-shaded frame in FP64: network residual over the network-free
    estimate, clamped, shading the march like the exact cone integrals; samples whose
    inputs leave [lo, hi] take the exact tracer.  Returns (image, margin, counts) with
    counts = (shaded samples, exact-tracer samples)."""
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    w = np.ascontiguousarray(weights, np.float64)
    assert w.size == lib().orc_mlp_nparams()
    lo = np.full(16, -1e300) if lo is None else np.ascontiguousarray(lo, np.float64)
    hi = np.full(16, 1e300) if hi is None else np.ascontiguousarray(hi, np.float64)
    out = np.zeros((H, W, 4), np.float64)
    margin = np.zeros((H, W, 2), np.float64)
    counts = np.zeros(2, np.int64)
    lib().orc_render_nn_f64(*va, tf.h, pyr.h, C.byref(prm), w.ctypes.data, lo.ctypes.data, hi.ctypes.data, int(bf16),
                            c.ctypes.data, W, H, step, et, out.ctypes.data, margin.ctypes.data, threads, stride,
                            counts.ctypes.data)
    return out, margin, counts


# ---------------------------------------------------------------------------- section tables (DOS)
class DosConeC(C.Structure):
    _fields_ = [("half_deg", C.c_double), ("max_packing", C.c_int), ("ui_weight", C.c_double),
                ("initial_step", C.c_double), ("covered", C.c_double), ("d_sigma", C.c_double),
                ("r_sigma", C.c_double)]


class DosC(C.Structure):
    _fields_ = [("apply_occ", C.c_int), ("apply_sdw", C.c_int), ("shadow_type", C.c_int), ("spot_cos", C.c_double),
                ("ka", C.c_double), ("kd", C.c_double), ("occ", DosConeC), ("sdw", DosConeC),
                ("lpos", C.c_double * 3), ("lfwd", C.c_double * 3), ("lup", C.c_double * 3),
                ("lright", C.c_double * 3)]


def _f32(x):
    return float(np.float32(x))


def diagonal_f32(grid):
    g = np.asarray(grid, np.float32)
    return np.float32(np.sqrt(np.float32(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])))


def dos_cone_spec(native_cone, covered_default) -> DosConeC:
    """A native.DosCone after ConeGaussianSampler's setter clamps
    (conegaussiansampler.cpp:57-172); covered_default is used when covered_dist <= 0."""
    c = DosConeC()
    c.half_deg = _f32(min(max(native_cone.half_deg, 0.5), 89.5))
    c.max_packing = min(max(native_cone.max_packing, 0), 2)
    c.ui_weight = _f32(native_cone.ui_weight)
    c.initial_step = _f32(native_cone.initial_step)
    cov = native_cone.covered_dist if native_cone.covered_dist > 0 else covered_default
    c.covered = _f32(max(_f32(cov), 10.0))
    c.d_sigma = _f32(min(max(native_cone.d_sigma, 1.0), 3.0))
    c.r_sigma = _f32(min(max(native_cone.r_sigma, 0.5), 3.0))
    return c


def dos_params(p, grid, light=None, cam=None) -> DosC:
    """From a native.DosParams, the volume's world size and the light (or, without one,
    the camera: a light at the eye along the view axes)."""
    d = DosC()
    d.apply_occ, d.apply_sdw, d.shadow_type = p.apply_occlusion, p.apply_shadow, p.shadow_type
    d.spot_cos = float(np.cos(p.spot_max_deg * np.pi / 180.0))
    d.ka, d.kd = _f32(p.ka), _f32(p.kd)
    diag = diagonal_f32(grid)
    d.occ = dos_cone_spec(p.occ, np.float32(diag * np.float32(0.5)))
    d.sdw = dos_cone_spec(p.sdw, np.float32(diag * np.float32(0.75)))
    if light is not None:
        d.lpos[:], d.lfwd[:], d.lup[:], d.lright[:] = (list(light.position), list(light.forward), list(light.up),
                                                       list(light.right))
    return d


def dos_sections(cone: DosConeC, sigma0=1.0, max_rows=1 << 16):
    rows = np.zeros((max_rows, 8))
    counts = np.zeros(3, np.int32)
    axes = np.zeros((10, 3))
    adj7 = C.c_double()
    n = lib().orc_dos_sections(C.byref(cone), sigma0, rows.ctypes.data, max_rows, counts.ctypes.data,
                               axes.ctypes.data, C.byref(adj7))
    if n < 0:
        raise ValueError("section table could not be built")
    return rows[:n], counts, axes, adj7.value


def dos_cone(pyr: OPyr, cone: DosConeC, pos, k, u, v, sigma0=1.0):
    a = [np.ascontiguousarray(x, np.float64) for x in (pos, k, u, v)]
    return lib().orc_dos_cone(pyr.h, C.byref(cone), sigma0, *[x.ctypes.data for x in a])


def render_dos(vol, tf: OTF, pyr: OPyr, prm: DosC, cam, W, H, sigma0=1.0, step=0.5, et=0.99, scale=(1, 1, 1),
               fp32=False, threads=0, stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    if fp32:
        out = np.zeros((H, W, 4), np.float32)
        lib().orc_render_dos_f32(*va, tf.h, pyr.h, C.byref(prm), sigma0, c.ctypes.data, W, H, step, et,
                                 out.ctypes.data, threads, stride)
        return out
    out = np.zeros((H, W, 4), np.float64)
    margin = np.zeros((H, W, 2), np.float64)
    lib().orc_render_dos_f64(*va, tf.h, pyr.h, C.byref(prm), sigma0, c.ctypes.data, W, H, step, et, out.ctypes.data,
                             margin.ctypes.data, threads, stride)
    return out, margin


# ---------------------------------------------------------------------------- cone ground truth
class GtC(C.Structure):
    _fields_ = [("apply_occ", C.c_int), ("apply_sdw", C.c_int), ("shadow_type", C.c_int), ("n_occ", C.c_int),
                ("n_sdw", C.c_int), ("occ_aperture_deg", C.c_double), ("sdw_aperture_deg", C.c_double),
                ("occ_dist", C.c_double), ("sdw_dist", C.c_double), ("gap", C.c_double), ("rstep", C.c_double),
                ("ka", C.c_double), ("kd", C.c_double), ("lpos", C.c_double * 3), ("lfwd", C.c_double * 3),
                ("lup", C.c_double * 3), ("lright", C.c_double * 3)]


def gt_params(p, grid, light=None) -> GtC:
    """From a native.GtParams; the shader's LightCamForward is minus the light's forward
    (crtgtrenderer.cpp:224-234)."""
    g = GtC()
    g.apply_occ, g.apply_sdw, g.shadow_type = p.apply_occlusion, p.apply_shadow, p.shadow_type
    g.n_occ, g.n_sdw = p.occ_rays, p.sdw_rays
    g.occ_aperture_deg, g.sdw_aperture_deg = _f32(p.occ_aperture_deg), _f32(p.sdw_aperture_deg)
    diag = diagonal_f32(grid)
    g.occ_dist = _f32(p.occ_max_dist) if p.occ_max_dist > 0 else float(np.float32(diag * np.float32(0.5)))
    g.sdw_dist = _f32(p.sdw_max_dist) if p.sdw_max_dist > 0 else float(np.float32(diag * np.float32(0.75)))
    g.gap, g.rstep, g.ka, g.kd = _f32(p.initial_gap), _f32(p.ray_step), _f32(p.ka), _f32(p.kd)
    if light is not None:
        g.lpos[:] = list(light.position)
        g.lfwd[:] = [-float(x) for x in light.forward]
        g.lup[:] = list(light.up)
        g.lright[:] = list(light.right)
    return g


def gt_rays(prm: GtC):
    occ = np.zeros((prm.n_occ, 3))
    sdw = np.zeros((prm.n_sdw, 3))
    lib().orc_gt_rays(C.byref(prm), occ.ctypes.data, sdw.ctypes.data)
    return occ, sdw


def gt_ray(vol, tf: OTF, pos, w, dist, gap=1.0, rstep=0.5, scale=(1, 1, 1)):
    vol, va = _vol_args(vol, scale)
    p = np.ascontiguousarray(pos, np.float64)
    ww = np.ascontiguousarray(w, np.float64)
    return lib().orc_gt_ray(*va, tf.h, p.ctypes.data, ww.ctypes.data, dist, gap, rstep)


def render_gt(vol, tf: OTF, prm: GtC, cam, W, H, step=0.5, et=0.99, scale=(1, 1, 1), fp32=False, threads=0,
              stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    if fp32:
        out = np.zeros((H, W, 4), np.float32)
        lib().orc_render_gt_f32(*va, tf.h, C.byref(prm), c.ctypes.data, W, H, step, et, out.ctypes.data, threads,
                                stride)
        return out
    out = np.zeros((H, W, 4), np.float64)
    margin = np.zeros((H, W, 2), np.float64)
    lib().orc_render_gt_f64(*va, tf.h, C.byref(prm), c.ctypes.data, W, H, step, et, out.ctypes.data,
                            margin.ctypes.data, threads, stride)
    return out, margin


# ---------------------------------------------------------------- light caches
def lc_ebs(shape, scale, sat_arr, prm: EbsC, res, threads=0):
    """EBS light cache at resolution res = (rx, ry, rz): (rz, ry, rx, 2) of (occlusion, shadow)."""
    nx, ny, nz = shape
    rx, ry, rz = res
    s64 = np.ascontiguousarray(sat_arr, dtype=np.float64)
    out = np.zeros((rz, ry, rx, 2))
    lib().orc_lc_ebs(nx, ny, nz, *map(float, scale), s64.ctypes.data, C.byref(prm), rx, ry, rz, out.ctypes.data,
                     threads)
    return out


def lc_dos(shape, scale, pyr: "OPyr", prm: "DosC", cam, W, H, res, sigma0=1.0, threads=0):
    """DOS light cache for the camera `cam` (its eye and up vector)."""
    nx, ny, nz = shape
    rx, ry, rz = res
    out = np.zeros((rz, ry, rx, 2))
    c = cam_array(cam)
    if lib().orc_lc_dos(nx, ny, nz, *map(float, scale), pyr.h, C.byref(prm), sigma0, c.ctypes.data, W, H, rx, ry, rz,
                        out.ctypes.data, threads) != 0:
        raise ValueError("cone sections do not converge")
    return out


def render_lc(vol, tf: OTF, cache, occ, sdw, ka, kd, cam, W, H, step=0.5, et=0.99, scale=(1, 1, 1), fp32=False,
              threads=0, stride=1):
    """The object-space march over a light cache (rz, ry, rx, 2)."""
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    cache = np.ascontiguousarray(cache, dtype=np.float64)
    rz, ry, rx, _ = cache.shape
    if fp32:
        out = np.zeros((H, W, 4), np.float32)
        lib().orc_render_lc_f32(*va, tf.h, cache.ctypes.data, rx, ry, rz, int(occ), int(sdw), float(ka), float(kd),
                                c.ctypes.data, W, H, step, et, out.ctypes.data, threads, stride)
        return out
    out = np.zeros((H, W, 4), np.float64)
    margin = np.zeros((H, W, 2), np.float64)
    lib().orc_render_lc_f64(*va, tf.h, cache.ctypes.data, rx, ry, rz, int(occ), int(sdw), float(ka), float(kd),
                            c.ctypes.data, W, H, step, et, out.ctypes.data, margin.ctypes.data, threads, stride)
    return out, margin


# ---------------------------------------------------------------- gradients, Blinn-Phong
class PhongC(C.Structure):
    _fields_ = [("type", C.c_int), ("ka", C.c_double), ("kd", C.c_double), ("ks", C.c_double),
                ("shininess", C.c_double), ("spec", C.c_double * 3), ("lpos", C.c_double * 3)]


def phong(kind=0, light=None, ka=0.5, kd=0.5, ks=0.8, shininess=30.0, spec=(1.0, 1.0, 1.0), lpos=None) -> PhongC:
    """Blinn-Phong constants (reference defaults); the light position from `light` or `lpos`."""
    p = PhongC()
    p.type, p.ka, p.kd, p.ks, p.shininess = kind, _f32(ka), _f32(kd), _f32(ks), _f32(shininess)
    p.spec[:] = [_f32(x) for x in spec]
    pos = lpos if lpos is not None else (light.position if light is not None else (0.0, 0.0, 0.0))
    p.lpos[:] = [float(np.float32(x)) for x in pos]
    return p


def gradient(vol, kind=0, threads=0):
    """The volume gradient (nz, ny, nx, 3): kind 0 Sobel-Feldman, 1 finite differences."""
    vol = np.ascontiguousarray(vol)
    nz, ny, nx = vol.shape
    out = np.zeros((nz, ny, nx, 3))
    lib().orc_gradient(nx, ny, nz, 0 if vol.dtype == np.uint8 else 1, vol.ctypes.data, kind, out.ctypes.data, threads)
    return out


def _phong_out(W, H):
    return np.zeros((H, W, 4), np.float64), np.zeros((H, W, 2), np.float64)


def render_rc1p_phong(vol, tf: OTF, ph: PhongC, cam, W, H, step=0.5, et=0.99, scale=(1, 1, 1), threads=0, stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    out, margin = _phong_out(W, H)
    lib().orc_render_rc1p_phong_f64(*va, tf.h, C.byref(ph), c.ctypes.data, W, H, step, et, out.ctypes.data,
                                    margin.ctypes.data, threads, stride)
    return out, margin


def render_ebs_phong(vol, tf: OTF, sat_arr, prm: EbsC, ph: PhongC, cam, W, H, step=0.5, et=0.99, scale=(1, 1, 1),
                     threads=0, stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    s64 = np.ascontiguousarray(sat_arr, dtype=np.float64)
    out, margin = _phong_out(W, H)
    lib().orc_render_ebs_phong_f64(*va, tf.h, s64.ctypes.data, C.byref(prm), C.byref(ph), c.ctypes.data, W, H, step,
                                   et, out.ctypes.data, margin.ctypes.data, threads, stride)
    return out, margin


def render_dos_phong(vol, tf: OTF, pyr, prm, ph: PhongC, cam, W, H, sigma0=1.0, step=0.5, et=0.99, scale=(1, 1, 1),
                     threads=0, stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    out, margin = _phong_out(W, H)
    lib().orc_render_dos_phong_f64(*va, tf.h, pyr.h, C.byref(prm), sigma0, C.byref(ph), c.ctypes.data, W, H, step, et,
                                   out.ctypes.data, margin.ctypes.data, threads, stride)
    return out, margin


def render_lc_phong(vol, tf: OTF, cache, occ, sdw, ka, kd, ph: PhongC, cam, W, H, step=0.5, et=0.99, scale=(1, 1, 1),
                    threads=0, stride=1):
    vol, va = _vol_args(vol, scale)
    c = cam_array(cam)
    cache = np.ascontiguousarray(cache, dtype=np.float64)
    rz, ry, rx, _ = cache.shape
    out, margin = _phong_out(W, H)
    lib().orc_render_lc_phong_f64(*va, tf.h, cache.ctypes.data, rx, ry, rz, int(occ), int(sdw), float(ka), float(kd),
                                  C.byref(ph), c.ctypes.data, W, H, step, et, out.ctypes.data, margin.ctypes.data,
                                  threads, stride)
    return out, margin


# ---------------------------------------------------------------- multi-scale output
def resample(frame, dw, dh, mode, kernel):
    """An (H, W, 4) frame resampled to (dh, dw, 4): mode 0 multisample, 1 downscale, 2 upscale."""
    src = np.ascontiguousarray(frame, dtype=np.float64)
    sh, sw, _ = src.shape
    out = np.zeros((dh, dw, 4))
    lib().orc_resample(src.ctypes.data, sw, sh, out.ctypes.data, dw, dh, mode, kernel)
    return out
