"""ctypes binding of the C ABI in ``include/cvr.h`` (``libcvr.so``, built in-tree).

The library is loaded from this package directory only; when it is missing the
import fails loudly (there is no CPU fallback for the HIP path).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libcvr.so"

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_STATE = -3
ERR_UNSUPPORTED = -4
ERR_NOMEM = -5

RC1P, EBS, CONE_GT, DOS, CONE_TRACE, CONE_NN = 0, 1, 2, 3, 4, 5
U8, U16 = 0, 1


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("center", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy_deg", C.c_float), ("aspect", C.c_float)]


class Light(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("forward", C.c_float * 3), ("up", C.c_float * 3),
                ("right", C.c_float * 3)]


GRADIENT_SOBEL, GRADIENT_FINITE_DIFF = 0, 1
(FILTER_BOX, FILTER_HAT, FILTER_CATMULL_ROM, FILTER_MITCHELL_NETRAVALI, FILTER_CARDINAL_BSPLINE3,
 FILTER_CARDINAL_OMOMS3) = range(6)
SCALE_MULTISAMPLE, SCALE_DOWNSCALE, SCALE_UPSCALE = 0, 1, 2
SINGLE_RAY_PER_PIXEL, MULTIPLE_RAYS_PER_PIXEL, DOWN_SCALING_RENDER, UP_SCALING_RENDER = 0, 1, 2, 3


class Rc1pParams(C.Structure):
    _fields_ = [("step_size", C.c_float), ("et_threshold", C.c_float), ("apply_gradient_shading", C.c_int),
                ("gradient_type", C.c_int), ("ka", C.c_float), ("kd", C.c_float), ("ks", C.c_float),
                ("shininess", C.c_float), ("specular", C.c_float * 3)]


class EbsParams(C.Structure):
    _fields_ = [("step_size", C.c_float), ("et_threshold", C.c_float), ("apply_occlusion", C.c_int),
                ("apply_shadow", C.c_int), ("ao_shells", C.c_int), ("ao_radius", C.c_float),
                ("sdw_angle_deg", C.c_float), ("sdw_interval", C.c_float), ("sdw_initial", C.c_float),
                ("sdw_weight", C.c_float), ("sdw_max_dist", C.c_float), ("shadow_type", C.c_int),
                ("ka", C.c_float), ("kd", C.c_float), ("light_cache", C.c_int), ("light_cache_res", C.c_int * 3),
                ("apply_phong", C.c_int), ("gradient_type", C.c_int), ("ks", C.c_float), ("shininess", C.c_float),
                ("specular", C.c_float * 3), ("sat_layout", C.c_int)]


class GtParams(C.Structure):
    _fields_ = [("step_size", C.c_float), ("et_threshold", C.c_float), ("apply_occlusion", C.c_int),
                ("apply_shadow", C.c_int), ("occ_rays", C.c_int), ("occ_aperture_deg", C.c_float),
                ("occ_max_dist", C.c_float), ("sdw_rays", C.c_int), ("sdw_aperture_deg", C.c_float),
                ("sdw_max_dist", C.c_float), ("shadow_type", C.c_int), ("initial_gap", C.c_float),
                ("ray_step", C.c_float), ("ka", C.c_float), ("kd", C.c_float)]


class DosCone(C.Structure):
    _fields_ = [("half_deg", C.c_float), ("max_packing", C.c_int), ("ui_weight", C.c_float),
                ("initial_step", C.c_float), ("covered_dist", C.c_float), ("d_sigma", C.c_float),
                ("r_sigma", C.c_float)]


class DosParams(C.Structure):
    _fields_ = [("step_size", C.c_float), ("et_threshold", C.c_float), ("apply_occlusion", C.c_int),
                ("apply_shadow", C.c_int), ("shadow_type", C.c_int), ("spot_max_deg", C.c_float),
                ("n_levels", C.c_int), ("base_res", C.c_int * 3), ("sigma0", C.c_float), ("occ", DosCone),
                ("sdw", DosCone), ("ka", C.c_float), ("kd", C.c_float), ("density_precision", C.c_int),
                ("light_cache", C.c_int), ("light_cache_res", C.c_int * 3), ("apply_phong", C.c_int),
                ("gradient_type", C.c_int), ("ks", C.c_float), ("shininess", C.c_float), ("specular", C.c_float * 3)]


class ConeParams(C.Structure):
    _fields_ = [("step_size", C.c_float), ("et_threshold", C.c_float), ("apply_occlusion", C.c_int),
                ("apply_shadow", C.c_int), ("n_levels", C.c_int), ("base_res", C.c_int * 3), ("sigma0", C.c_float),
                ("n_steps", C.c_int), ("n_occ_cones", C.c_int), ("occ_half_deg", C.c_float),
                ("ring_deg", C.c_float), ("sdw_half_deg", C.c_float), ("t0", C.c_float),
                ("occ_max_dist", C.c_float), ("sdw_max_dist", C.c_float), ("r_sigma", C.c_float),
                ("ka", C.c_float), ("kd", C.c_float), ("density_precision", C.c_int)]


class NnParams(C.Structure):
    _fields_ = [("cone", ConeParams), ("n_train", C.c_int), ("n_val", C.c_int), ("epochs", C.c_int),
                ("batch", C.c_int), ("lr", C.c_float), ("rmse_gate", C.c_float), ("ood_margin", C.c_float),
                ("force_exact", C.c_int), ("seed", C.c_uint64), ("tile_w", C.c_int), ("shade_batch", C.c_float),
                ("span", C.c_int)]


class NnReport(C.Structure):
    _fields_ = [("train_loss", C.c_double), ("val_rmse_occ", C.c_double), ("val_rmse_sdw", C.c_double),
                ("base_rmse_occ", C.c_double), ("base_rmse_sdw", C.c_double), ("steps", C.c_int),
                ("exact_fallback", C.c_int)]


class Job(C.Structure):
    _fields_ = [("volume", C.c_int), ("camera", Camera), ("light", Light)]


class BatchTuning(C.Structure):
    _fields_ = [("tile_w", C.c_int), ("shade_batch", C.c_float), ("queues", C.c_int),
                ("workgroups_per_cu", C.c_int), ("span", C.c_int), ("ms", C.c_double), ("step_scale", C.c_float),
                ("max_step_rmse", C.c_float), ("step_rmse", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("shaded_samples", C.c_uint64), ("fallback_samples", C.c_uint64),
                ("last_ms", C.c_double), ("box_sums", C.c_uint64), ("sat_scan_ms", C.c_double),
                ("sat_blocks_ms", C.c_double)]


PARAMS = {RC1P: Rc1pParams, EBS: EbsParams, CONE_GT: GtParams, DOS: DosParams, CONE_TRACE: ConeParams,
          CONE_NN: NnParams}

# every symbol include/cvr.h and include/cvr_synth.h declare
EXPORTS = [
    "cvr_last_error", "cvr_technique_name", "cvr_create", "cvr_destroy", "cvr_set_volume",
    "cvr_set_transfer_function", "cvr_default_params", "cvr_set_params", "cvr_get_params", "cvr_prepare",
    "cvr_render", "cvr_render_noalloc", "cvr_render_multiscale", "cvr_probe_gather", "cvr_get_sat", "cvr_get_light_cache", "cvr_get_gradient", "cvr_get_stats", "cvr_set_collect_stats",
    "cvr_multiscale_size", "cvr_resample_frame",
    "cvr_pyramid_info", "cvr_get_pyramid_level", "cvr_cone_eval", "cvr_nn_param_count", "cvr_nn_train",
    "cvr_nn_get_weights", "cvr_nn_set_weights", "cvr_nn_eval", "cvr_nn_gradient", "cvr_nn_get_report",
    "cvr_nn_checkpoint_size", "cvr_nn_save", "cvr_nn_load", "cvr_dos_sections", "cvr_gt_rays",
    "cvr_batch_create", "cvr_batch_destroy", "cvr_batch_add_volume", "cvr_batch_set_jobs",
    "cvr_batch_set_tuning", "cvr_batch_get_tuning", "cvr_batch_autotune", "cvr_batch_render",
    "cvr_batch_get_stats",
    "cvr_synth_blobs", "cvr_synth_noise",
]

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m cpp_volume_rendering_amd.build` "
                          "(there is no CPU fallback for the HIP path)")
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7) and loads it under a different file name, so a process that
    # loads libcvr first and torch later would hold two runtimes, and streams or
    # device pointers would not cross between them.  Loading torch first lets libcvr
    # bind to the runtime already present by soname.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    vp, i, f, d, sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_size_t
    sig = {
        "cvr_last_error": (C.c_char_p, []),
        "cvr_technique_name": (C.c_char_p, [i]),
        "cvr_create": (i, [i, i, C.POINTER(vp)]),
        "cvr_destroy": (None, [vp]),
        "cvr_set_volume": (i, [vp, i, i, i, i, f, f, f, vp, i]),
        "cvr_set_transfer_function": (i, [vp, i, i, i, vp, i, vp]),
        "cvr_default_params": (i, [i, vp, sz]),
        "cvr_set_params": (i, [vp, vp, sz]),
        "cvr_get_params": (i, [vp, vp, sz]),
        "cvr_prepare": (i, [vp, vp]),
        "cvr_render": (i, [vp, C.POINTER(Camera), C.POINTER(Light), i, i, vp, i, vp]),
        "cvr_render_noalloc": (i, [vp, C.POINTER(Camera), C.POINTER(Light), i, i, vp, vp]),
        "cvr_probe_gather": (i, [i, C.c_size_t, i, C.POINTER(C.c_double), C.POINTER(C.c_double), vp]),
        "cvr_render_multiscale": (i, [vp, C.POINTER(Camera), C.POINTER(Light), i, i, i, i, i, i, vp, i, vp]),
        "cvr_get_sat": (i, [vp, vp, sz]),
        "cvr_get_light_cache": (i, [vp, vp, sz, vp]),
        "cvr_get_gradient": (i, [vp, i, vp, sz]),
        "cvr_multiscale_size": (i, [i, i, i, i, vp, vp]),
        "cvr_resample_frame": (i, [i, vp, i, i, vp, i, i, i, i, vp, vp]),
        "cvr_get_stats": (i, [vp, C.POINTER(Stats)]),
        "cvr_set_collect_stats": (i, [vp, i]),
        "cvr_pyramid_info": (i, [vp, vp, vp, i]),
        "cvr_get_pyramid_level": (i, [vp, i, vp, sz]),
        "cvr_cone_eval": (i, [vp, i, vp, vp, vp, vp, vp, vp]),
        "cvr_nn_param_count": (i, []),
        "cvr_nn_train": (i, [vp, vp]),
        "cvr_nn_get_weights": (i, [vp, vp, sz]),
        "cvr_nn_set_weights": (i, [vp, vp, sz]),
        "cvr_nn_eval": (i, [vp, i, vp, vp, i]),
        "cvr_nn_gradient": (i, [vp, i, vp, vp, vp, vp]),
        "cvr_nn_get_report": (i, [vp, C.POINTER(NnReport)]),
        "cvr_nn_checkpoint_size": (sz, []),
        "cvr_nn_save": (i, [vp, vp, sz]),
        "cvr_nn_load": (i, [vp, vp, sz]),
        "cvr_dos_sections": (i, [vp, i, vp, i, vp, vp, vp, vp]),
        "cvr_gt_rays": (i, [vp, i, vp, i, vp]),
        "cvr_batch_create": (i, [i, i, i, C.POINTER(vp)]),
        "cvr_batch_destroy": (None, [vp]),
        "cvr_batch_add_volume": (i, [vp, vp, C.POINTER(i)]),
        "cvr_batch_set_jobs": (i, [vp, i, vp]),
        "cvr_batch_set_tuning": (i, [vp, C.POINTER(BatchTuning)]),
        "cvr_batch_get_tuning": (i, [vp, C.POINTER(BatchTuning)]),
        "cvr_batch_autotune": (i, [vp, vp, i, vp, C.POINTER(BatchTuning)]),
        "cvr_batch_render": (i, [vp, vp, vp]),
        "cvr_batch_get_stats": (i, [vp, C.POINTER(Stats)]),
        "cvr_synth_blobs": (i, [i, i, i, i, i, d, i, C.c_uint64, vp]),
        "cvr_synth_noise": (i, [i, i, i, i, i, C.c_uint64, i, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class CvrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cvr error {code}: {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != OK:
        raise CvrError(rc, (lib().cvr_last_error() or b"").decode())


def _ptr(a) -> int:
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())  # torch tensor


def _out_arg(out, shape, device=None):
    """Checks an output buffer before its address goes to the C ABI: float32,
    C-contiguous, exactly `shape`; a numpy array is host memory, a torch tensor must live
    on a GPU (on `device` when given).  Returns (pointer, is_device)."""
    shape = tuple(int(x) for x in shape)
    if isinstance(out, np.ndarray):
        if out.dtype != np.float32 or not out.flags["C_CONTIGUOUS"] or out.shape != shape:
            raise CvrError(ERR_INVALID, f"out must be a C-contiguous float32 array of shape {shape}")
        return out.ctypes.data, False
    try:
        import torch
    except ImportError as e:  # pragma: no cover - torch is how device buffers are passed
        raise CvrError(ERR_INVALID, "out must be a numpy array or a torch tensor") from e
    if not isinstance(out, torch.Tensor):
        raise CvrError(ERR_INVALID, "out must be a numpy array or a torch tensor")
    if out.dtype != torch.float32 or not out.is_contiguous() or tuple(out.shape) != shape:
        raise CvrError(ERR_INVALID, f"out must be a contiguous float32 tensor of shape {shape}")
    if not out.is_cuda:
        raise CvrError(ERR_INVALID, "out tensor must be on a GPU (use a numpy array for host output)")
    if device is not None and out.device.index != device:
        raise CvrError(ERR_INVALID, f"out tensor is on cuda:{out.device.index}, the handle on device {device}")
    return int(out.data_ptr()), True


def _volume_dtype(dt_name: str) -> int:
    if dt_name in ("uint8", "torch.uint8"):
        return U8
    if dt_name in ("uint16", "torch.uint16"):
        return U16
    raise CvrError(ERR_INVALID, f"volume dtype must be uint8 or uint16, not {dt_name}")


def synth_blobs(shape, dtype=U8, n_blobs=8, seed=1, noise_weight=0.0, noise_octaves=4) -> np.ndarray:
    nx, ny, nz = shape
    out = np.empty((nz, ny, nx), dtype=np.uint8 if dtype == U8 else np.uint16)
    check(lib().cvr_synth_blobs(nx, ny, nz, dtype, n_blobs, noise_weight, noise_octaves, seed, out.ctypes.data))
    return out


def synth_noise(shape, dtype=U16, octaves=4, seed=2, bands=None) -> np.ndarray:
    nx, ny, nz = shape
    out = np.empty((nz, ny, nx), dtype=np.uint8 if dtype == U8 else np.uint16)
    bv = None if bands is None else np.asarray(bands, dtype=np.float64)
    check(lib().cvr_synth_noise(nx, ny, nz, dtype, octaves, seed, 0 if bv is None else len(bv),
                                None if bv is None else bv.ctypes.data, out.ctypes.data))
    return out


class Renderer:
    """One technique instance (a cvr_handle)."""

    def __init__(self, technique: int, device: int = 0):
        self.technique = technique
        self.device = device
        h = C.c_void_p()
        check(lib().cvr_create(technique, device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().cvr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_volume(self, vol, scale=(1.0, 1.0, 1.0)):
        """vol: numpy (nz, ny, nx) uint8/uint16 on the host, or a torch tensor on the device."""
        if vol.ndim != 3:
            raise CvrError(ERR_INVALID, "volume must be 3-dimensional (nz, ny, nx)")
        nz, ny, nx = vol.shape
        if isinstance(vol, np.ndarray):
            dt = _volume_dtype(str(vol.dtype))
            vol = np.ascontiguousarray(vol)
            check(lib().cvr_set_volume(self.h, nx, ny, nz, dt, *scale, vol.ctypes.data, 0))
        else:
            dt = _volume_dtype(str(vol.dtype))
            if not vol.is_cuda or not vol.is_contiguous():
                raise CvrError(ERR_INVALID, "a torch volume must be a contiguous GPU tensor")
            check(lib().cvr_set_volume(self.h, nx, ny, nz, dt, *scale, vol.data_ptr(), 1))

    def set_transfer_function(self, rgb_iso, alpha_iso, max_density=255, extinction_input=False):
        rgb = np.ascontiguousarray(rgb_iso, dtype=np.float64).reshape(-1, 4)
        a = np.ascontiguousarray(alpha_iso, dtype=np.float64).reshape(-1, 2)
        check(lib().cvr_set_transfer_function(self.h, max_density, int(extinction_input), len(rgb),
                                              rgb.ctypes.data, len(a), a.ctypes.data))

    def default_params(self):
        p = PARAMS[self.technique]()
        check(lib().cvr_default_params(self.technique, C.byref(p), C.sizeof(p)))
        return p

    def params(self):
        p = PARAMS[self.technique]()
        check(lib().cvr_get_params(self.h, C.byref(p), C.sizeof(p)))
        return p

    def set_params(self, p=None, **kw):
        p = p if p is not None else self.params()
        for k, v in kw.items():
            setattr(p, k, v)
        check(lib().cvr_set_params(self.h, C.byref(p), C.sizeof(p)))
        return p

    def prepare(self, stream=None):
        check(lib().cvr_prepare(self.h, stream))

    def render(self, cam: Camera, width: int, height: int, light: Light | None = None, out=None, stream=None):
        """Render to a new host array (H, W, 4) float32, or into `out` (numpy or device torch tensor)."""
        if out is None:
            out = np.zeros((height, width, 4), dtype=np.float32)
        ptr, is_dev = _out_arg(out, (height, width, 4), self.device)
        check(lib().cvr_render(self.h, C.byref(cam), C.byref(light) if light is not None else None, width, height,
                               ptr, int(is_dev), stream))
        return out

    def render_multiscale(self, cam: Camera, width: int, height: int, mode: int, light: Light | None = None,
                          mult=(2, 2), kernel=None, out=None, stream=None):
        """cvr_render_multiscale: render at the mode's multi-scale frame size and resample
        to (height, width, 4) (a new host array, or `out`)."""
        if out is None:
            out = np.zeros((height, width, 4), dtype=np.float32)
        ptr, is_dev = _out_arg(out, (height, width, 4), self.device)
        k = FILTER_HAT if kernel is None else kernel
        check(lib().cvr_render_multiscale(self.h, C.byref(cam), C.byref(light) if light is not None else None, width,
                                          height, mode, mult[0], mult[1], k, ptr, int(is_dev), stream))
        return out

    def render_noalloc(self, cam: Camera, width: int, height: int, out, light: Light | None = None, stream=None):
        """cvr_render_noalloc: render a prepared CVR_CONE_NN handle into the device tensor
        `out` ((H, W, 4) float32) without any device allocation."""
        ptr, is_dev = _out_arg(out, (height, width, 4), self.device)
        if not is_dev:
            raise CvrError(ERR_INVALID, "cvr_render_noalloc writes device memory only")
        check(lib().cvr_render_noalloc(self.h, C.byref(cam), C.byref(light) if light is not None else None, width,
                                       height, ptr, stream))
        return out

    def gradient(self, shape, kind=GRADIENT_SOBEL) -> np.ndarray:
        """The volume gradient used by Blinn-Phong shading: (nz, ny, nx, 3) float32."""
        nx, ny, nz = shape
        out = np.empty((nz, ny, nx, 3), dtype=np.float32)
        check(lib().cvr_get_gradient(self.h, kind, out.ctypes.data, out.size))
        return out

    def light_cache(self) -> np.ndarray:
        """The light cache of the last render: (rz, ry, rx, 2) of (occlusion, shadow)."""
        res = np.zeros(3, np.int32)
        check(lib().cvr_get_light_cache(self.h, None, 0, res.ctypes.data))
        rx, ry, rz = (int(v) for v in res)
        out = np.empty((rz, ry, rx, 2), dtype=np.float32)
        check(lib().cvr_get_light_cache(self.h, out.ctypes.data, out.size, res.ctypes.data))
        return out

    def get_sat(self, shape) -> np.ndarray:
        nx, ny, nz = shape
        out = np.empty((nz + 2, ny + 2, nx + 2), dtype=np.float64)
        check(lib().cvr_get_sat(self.h, out.ctypes.data, out.size))
        return out

    # ------------------------------------------------------------ cone tracer / surrogate
    def pyramid(self):
        """List of the pyramid's extinction levels as (nz, ny, nx) float32 arrays."""
        n = C.c_int()
        dims = np.zeros(3 * 16, np.int32)
        check(lib().cvr_pyramid_info(self.h, C.byref(n), dims.ctypes.data, 16))
        out = []
        for l in range(n.value):
            nx, ny, nz = dims[3 * l:3 * l + 3]
            a = np.empty((nz, ny, nx), np.float32)
            check(lib().cvr_get_pyramid_level(self.h, l, a.ctypes.data, a.size))
            out.append(a)
        return out

    def cone_eval(self, pos, dirs, lights):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 3)
        n = len(pos)
        occ, sdw, feat = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros((n, 16), np.float32)
        check(lib().cvr_cone_eval(self.h, n, pos.ctypes.data, dirs.ctypes.data, lights.ctypes.data,
                                  occ.ctypes.data, sdw.ctypes.data, feat.ctypes.data))
        return occ, sdw, feat

    def dos_sections(self, which: int):
        """Section table of the occlusion (0) or shadow (1) cone: (rows (n, 4) float32,
        counts (3,), axes (10, 3), adj7)."""
        n = C.c_int()
        check(lib().cvr_dos_sections(self.h, which, None, 0, C.byref(n), None, None, None))
        rows = np.zeros((n.value, 4), np.float32)
        counts = np.zeros(3, np.int32)
        axes = np.zeros((10, 3), np.float32)
        adj7 = C.c_float()
        check(lib().cvr_dos_sections(self.h, which, rows.ctypes.data, n.value, C.byref(n), counts.ctypes.data,
                                     axes.ctypes.data, C.byref(adj7)))
        return rows, counts, axes, adj7.value

    def gt_rays(self, which: int) -> np.ndarray:
        n = C.c_int()
        check(lib().cvr_gt_rays(self.h, which, None, 0, C.byref(n)))
        out = np.zeros((n.value, 3), np.float32)
        check(lib().cvr_gt_rays(self.h, which, out.ctypes.data, n.value, C.byref(n)))
        return out

    def nn_train(self, stream=None):
        check(lib().cvr_nn_train(self.h, stream))

    def nn_weights(self):
        w = np.zeros(lib().cvr_nn_param_count(), np.float32)
        check(lib().cvr_nn_get_weights(self.h, w.ctypes.data, w.size))
        return w

    def nn_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        check(lib().cvr_nn_set_weights(self.h, w.ctypes.data, w.size))

    def nn_eval(self, X, path=1):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 16)
        Y = np.zeros((len(X), 2), np.float32)
        check(lib().cvr_nn_eval(self.h, len(X), X.ctypes.data, Y.ctypes.data, path))
        return Y

    def nn_gradient(self, X, T):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 16)
        T = np.ascontiguousarray(T, np.float32).reshape(-1, 2)
        g = np.zeros(lib().cvr_nn_param_count(), np.float32)
        loss = C.c_float()
        check(lib().cvr_nn_gradient(self.h, len(X), X.ctypes.data, T.ctypes.data, g.ctypes.data, C.byref(loss)))
        return loss.value, g

    def nn_save(self) -> bytes:
        n = lib().cvr_nn_checkpoint_size()
        buf = C.create_string_buffer(n)
        check(lib().cvr_nn_save(self.h, buf, n))
        return buf.raw

    def nn_load(self, blob: bytes):
        check(lib().cvr_nn_load(self.h, blob, len(blob)))

    def nn_range(self):
        """The training-set input range (lo[16], hi[16]) stored in the checkpoint: the
        renderer's exact tracer takes every sample with an input outside
        [lo - ood_margin, hi + ood_margin]."""
        blob = self.nn_save()
        o = 16 + 3 * lib().cvr_nn_param_count() * 4
        r = np.frombuffer(blob, np.float32, 32, o)
        return r[:16].copy(), r[16:].copy()

    def nn_report(self) -> NnReport:
        r = NnReport()
        check(lib().cvr_nn_get_report(self.h, C.byref(r)))
        return r

    def collect_stats(self, on=True):
        check(lib().cvr_set_collect_stats(self.h, int(on)))

    def stats(self) -> Stats:
        s = Stats()
        check(lib().cvr_get_stats(self.h, C.byref(s)))
        return s


class Batch:
    """Batched surrogate rendering (cvr_batch_*): CVR_CONE_NN renderers x jobs of one size."""

    def __init__(self, width: int, height: int, device: int = 0):
        h = C.c_void_p()
        check(lib().cvr_batch_create(device, width, height, C.byref(h)))
        self.h, self.width, self.height, self.device = h, width, height, device
        self._refs = []  # the renderers must outlive the batch
        self.n_jobs = 0

    def close(self):
        if self.h:
            lib().cvr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_volume(self, r: "Renderer") -> int:
        idx = C.c_int()
        check(lib().cvr_batch_add_volume(self.h, r.h, C.byref(idx)))
        self._refs.append(r)
        return idx.value

    def set_jobs(self, jobs):
        """jobs: list of (volume index, Camera, Light)."""
        arr = (Job * len(jobs))()
        for i, (v, cam, light) in enumerate(jobs):
            arr[i].volume, arr[i].camera, arr[i].light = v, cam, light
        check(lib().cvr_batch_set_jobs(self.h, len(jobs), arr))
        self.n_jobs = len(jobs)

    def tuning(self) -> BatchTuning:
        t = BatchTuning()
        check(lib().cvr_batch_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        names = {f[0] for f in BatchTuning._fields_}
        for k, v in kw.items():
            if k not in names:
                raise ValueError(f"unknown batch tuning field {k!r}")
            setattr(t, k, v)
        check(lib().cvr_batch_set_tuning(self.h, C.byref(t)))
        return t

    def _dev_out(self, out):
        # with no jobs set the C ABI reports the state error, whatever the leading size
        n = self.n_jobs if self.n_jobs else max(1, int(out.shape[0]))
        ptr, is_dev = _out_arg(out, (n, self.height, self.width, 4), self.device)
        if not is_dev:
            raise CvrError(ERR_INVALID, "batch output must be a GPU tensor")
        return ptr

    def autotune(self, out, stride=1, stream=None, max_step_rmse=None) -> BatchTuning:
        """Tunes the work shape (and, with max_step_rmse > 0, the step schedule first) on
        every stride-th job; see cvr_batch_autotune."""
        if max_step_rmse is not None:
            self.set_tuning(max_step_rmse=max_step_rmse)
        t = BatchTuning()
        check(lib().cvr_batch_autotune(self.h, self._dev_out(out), stride, stream, C.byref(t)))
        return t

    def render(self, out, stream=None):
        """out: device tensor (n_jobs, height, width, 4) float32."""
        check(lib().cvr_batch_render(self.h, self._dev_out(out), stream))
        return out

    def stats(self) -> Stats:
        s = Stats()
        check(lib().cvr_batch_get_stats(self.h, C.byref(s)))
        return s


def probe_gather(working_set_bytes: int, reps: int = 3, device: int = 0, stream=None):
    """GB/s of independent 32-byte record reads from a working set of the given size."""
    g, n = C.c_double(), C.c_double()
    check(lib().cvr_probe_gather(device, int(working_set_bytes), reps, C.byref(g), C.byref(n), stream))
    return g.value


def multiscale_size(screen_w: int, screen_h: int, mw: int, mh: int):
    """Frame size for a screen and multipliers (m > 0 multiplies, m < 0 divides)."""
    w, h = C.c_int(), C.c_int()
    check(lib().cvr_multiscale_size(screen_w, screen_h, mw, mh, C.byref(w), C.byref(h)))
    return w.value, h.value


def resample_frame(src, dst, mode=SCALE_DOWNSCALE, kernel=FILTER_HAT, work=None, stream=None, device=0):
    """Resample the device frame `src` (torch (sh, sw, 4) float32) into `dst` ((dh, dw, 4))."""
    sh, sw = int(src.shape[0]), int(src.shape[1])
    dh, dw = int(dst.shape[0]), int(dst.shape[1])
    check(lib().cvr_resample_frame(device, _ptr(src), sw, sh, _ptr(dst), dw, dh, mode, kernel,
                                   None if work is None else _ptr(work), stream))
    return dst
