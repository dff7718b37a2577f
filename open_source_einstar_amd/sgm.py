"""Python binding of the gfx950 SGM library through its C ABI (include/sgm_hip.h).

`HipStereo` mirrors the reference's GPU matcher, depth_metal::MetalStereo
(libs/depth_metal/include/einstar/depth_metal/metal_stereo.hpp:67-107): ``create`` validates the
parameters and sizes the device buffers, ``compute`` runs depth::sgm_disparity
(libs/depth/src/stereo.cpp:361-463) on a rectified pair, ``last_timings`` reports per-stage GPU
time. Errors raise `SgmError` carrying the einstar::Errc code (libs/core/include/einstar/core/error.hpp:8-19).

There is no CPU fallback: importing works anywhere, but every call needs ``libsgm_hip.so`` and a
HIP device, and fails loudly without them.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libsgm_hip.so")

INVALID_DISPARITY = np.float32(-1.0e9)
NO_MATCH = np.int32(-(2**31))
STATUS_WINDOW, STATUS_UNIQUE, STATUS_LR, STATUS_VALID = 1, 2, 4, 7
OPT_COST_VOLUME = 1
OPT_PATH_LINES = 2
OPT_PATH_GROUP = 3
OPT_RECOMPUTE_DIRS = 4
OPT_VIEW_BASE_HIGH = 5
OPT_ROW_THREADS = 6

ERRC = {0: "ok", 1: "not_found", 2: "io", 3: "timeout", 4: "protocol", 5: "invalid_argument", 6: "blocked",
        7: "unsupported", 8: "busy", 9: "disconnected"}


class SgmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERRC.get(code, code)}: {message}")
        self.code = code
        self.errc = ERRC.get(code, str(code))


class CSgmParams(C.Structure):
    _fields_ = [
        ("min_disparity", C.c_int32), ("num_disparities", C.c_int32), ("census_radius_x", C.c_int32),
        ("census_radius_y", C.c_int32), ("p1", C.c_int32), ("p2", C.c_int32), ("adaptive_p2", C.c_int32),
        ("eight_paths", C.c_int32), ("slant_steps", C.c_int32), ("uniqueness", C.c_float),
        ("lr_max_diff", C.c_int32), ("subpixel", C.c_int32),
    ]


OUTPUT_FIELDS = ["disparity", "disparity_median", "confidence", "disparity_int", "status", "right_disparity",
                 "best_cost", "second_cost", "sum_volume", "cost_volume", "census_left", "census_right"]
OUTPUT_DTYPES = {"disparity": np.float32, "disparity_median": np.float32, "confidence": np.float32,
                 "disparity_int": np.int32, "status": np.uint8, "right_disparity": np.int32,
                 "best_cost": np.uint16, "second_cost": np.uint16, "sum_volume": np.uint16,
                 "cost_volume": np.uint16, "census_left": np.uint64, "census_right": np.uint64}
VOLUME_FIELDS = ("sum_volume", "cost_volume")


class CSgmOutputs(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in OUTPUT_FIELDS]


class CSgmTimings(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("census_ms", "paths_ms", "rows_ms", "median_ms", "gpu_ms", "total_ms",
                                          "rectify_ms", "refine_ms", "filter_ms", "points_ms")]


class CSgmStageTotals(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("census_ms", "paths_ms", "rows_ms", "median_ms", "gpu_ms")] + \
               [("count", C.c_int64)] + [(f, C.c_double) for f in ("rectify_ms", "refine_ms", "filter_ms", "points_ms")]


class CSgmServerConfig(C.Structure):
    _fields_ = [("max_width", C.c_int32), ("max_height", C.c_int32), ("max_disparities", C.c_int32),
                ("slots", C.c_int32), ("hbm_budget_bytes", C.c_int64), ("compute_streams", C.c_int32),
                ("path_group", C.c_int32)]


class CSgmRequest(C.Structure):
    _fields_ = [("left", C.c_void_p), ("right", C.c_void_p), ("stride", C.c_int64), ("width", C.c_int32),
                ("height", C.c_int32), ("min_disparity", C.c_int32), ("num_disparities", C.c_int32),
                ("out", CSgmOutputs), ("user_tag", C.c_uint64), ("rectification", C.c_int32),
                ("borrow_inputs", C.c_int32)]


class CSgmCompletion(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("user_tag", C.c_uint64), ("status", C.c_int32),
                ("t_submit_ns", C.c_int64), ("t_admit_ns", C.c_int64), ("t_done_ns", C.c_int64),
                ("gpu_ms", C.c_double)]


class CSgmServerStats(C.Structure):
    _fields_ = [("queued", C.c_int32), ("in_flight", C.c_int32), ("completed", C.c_int64), ("failed", C.c_int64),
                ("arena_bytes", C.c_int64), ("arena_free_bytes", C.c_int64)]


class CSgmBatchTuning(C.Structure):
    _fields_ = [("cost_volume", C.c_int32), ("path_lines", C.c_int32), ("branches", C.c_int32),
                ("path_group", C.c_int32), ("sweep_band", C.c_int32), ("sweep_tile", C.c_int32)]


class CSgmBatchPlan(C.Structure):
    _fields_ = [("cost_volume", C.c_int32), ("path_lines", C.c_int32), ("branches", C.c_int32),
                ("candidates_timed", C.c_int32), ("best_ms", C.c_double), ("worst_ms", C.c_double),
                ("from_cache", C.c_int32), ("path_group", C.c_int32), ("sweep_band", C.c_int32),
                ("sweep_tile", C.c_int32)]


class CSgmPair(C.Structure):
    _fields_ = [("left", C.c_void_p), ("right", C.c_void_p), ("pitch", C.c_int64), ("width", C.c_int32),
                ("height", C.c_int32), ("min_disparity", C.c_int32), ("num_disparities", C.c_int32),
                ("out", CSgmOutputs)]


@dataclass
class SgmParams:
    """depth::SgmParams (libs/depth/include/einstar/depth/stereo.hpp:26-40) + StereoParams::subpixel."""

    min_disparity: int = 0
    num_disparities: int = 64
    census_radius_x: int = 2
    census_radius_y: int = 2
    p1: int = 8
    p2: int = 96
    adaptive_p2: bool = True
    eight_paths: bool = False
    slant_steps: bool = True
    uniqueness: float = 0.9
    lr_max_diff: int = 1
    subpixel: bool = True

    def to_c(self) -> CSgmParams:
        return CSgmParams(self.min_disparity, self.num_disparities, self.census_radius_x, self.census_radius_y,
                          self.p1, self.p2, int(self.adaptive_p2), int(self.eight_paths), int(self.slant_steps),
                          float(self.uniqueness), self.lr_max_diff, int(self.subpixel))


_STEREO_FIELDS = [
    ("pyramid_levels", C.c_int32), ("refine_search_radius", C.c_int32), ("refine_zncc_radius", C.c_int32),
    ("refine_min_zncc", C.c_float), ("refine_reject_edge_winners", C.c_int32), ("detail_enabled", C.c_int32),
    ("detail_radius", C.c_int32), ("detail_search", C.c_float), ("detail_step", C.c_float),
    ("detail_min_zncc", C.c_float), ("filter_occlusion", C.c_int32), ("filter_occlusion_tolerance", C.c_float),
    ("filter_spikes", C.c_int32), ("filter_spike_threshold", C.c_float), ("speckle_max_region_size", C.c_int32),
    ("speckle_max_diff", C.c_float),
]


class CSgmStereoParams(C.Structure):
    _fields_ = [("sgm", CSgmParams)] + _STEREO_FIELDS


@dataclass
class StereoParams:
    """depth::StereoParams (libs/depth/include/einstar/depth/stereo.hpp:42-89) minus sgm/subpixel, with
    the reference defaults: two pyramid levels and every stage on."""

    pyramid_levels: int = 2
    refine_search_radius: int = 2
    refine_zncc_radius: int = 3
    refine_min_zncc: float = 0.5
    refine_reject_edge_winners: bool = True
    detail_enabled: bool = True
    detail_radius: int = 5
    detail_search: float = 1.0
    detail_step: float = 0.5
    detail_min_zncc: float = 0.5
    filter_occlusion: bool = True
    filter_occlusion_tolerance: float = 1.0
    filter_spikes: bool = True
    filter_spike_threshold: float = 1.0
    speckle_max_region_size: int = 200
    speckle_max_diff: float = 1.0

    def to_c(self, sgm: "SgmParams") -> CSgmStereoParams:
        vals = [float(getattr(self, f)) if t is C.c_float else int(getattr(self, f)) for f, t in _STEREO_FIELDS]
        return CSgmStereoParams(sgm.to_c(), *vals)


class CSgmRectifiedGeometry(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("f", "cx", "cy", "baseline", "cx_offset")]


class CSgmPointParams(C.Structure):
    _fields_ = [(f, C.c_float) for f in ("min_depth", "max_depth", "max_depth_jump", "edge_weight",
                                         "weight_reference_depth")]


class CSgmFrame(C.Structure):
    _fields_ = [("points_xyzw", C.c_void_p), ("normals_xyzw", C.c_void_p), ("weights", C.c_void_p)]


@dataclass
class RectifiedGeometry:
    """depth::RectifiedGeometry (libs/depth/include/einstar/depth/point_image.hpp:12-26)."""

    f: float
    cx: float
    cy: float
    baseline: float
    cx_offset: float = 0.0

    def to_c(self) -> CSgmRectifiedGeometry:
        return CSgmRectifiedGeometry(self.f, self.cx, self.cy, self.baseline, self.cx_offset)


@dataclass
class PointParams:
    """depth::PointImageParams (point_image.hpp:43-49) with the reference defaults."""

    min_depth: float = 150.0
    max_depth: float = 700.0
    max_depth_jump: float = 4.0
    edge_weight: float = 0.25
    weight_reference_depth: float = 0.0

    def to_c(self) -> CSgmPointParams:
        return CSgmPointParams(self.min_depth, self.max_depth, self.max_depth_jump, self.edge_weight,
                               self.weight_reference_depth)


_lib = None


def lib():
    """Loads libsgm_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m open_source_einstar_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i32, P = C.c_void_p, C.c_int, C.POINTER
        L.sgm_params_default.argtypes = [P(CSgmParams)]
        L.sgm_create.argtypes = [P(CSgmParams), i32, i32, P(vp)]
        L.sgm_destroy.argtypes = [vp]
        L.sgm_last_error.argtypes = [vp]
        L.sgm_last_error.restype = C.c_char_p
        L.sgm_compute.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, P(CSgmOutputs)]
        L.sgm_compute_device.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, P(CSgmOutputs), vp]
        L.sgm_set_rectification.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, i32]
        L.sgm_compute_raw.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, P(CSgmOutputs), vp, vp]
        L.sgm_compute_raw_device.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, P(CSgmOutputs), vp, vp, vp]
        SP = P(CSgmStereoParams)
        L.sgm_stereo_params_default.argtypes = [SP]
        L.sgm_create_stereo.argtypes = [SP, i32, i32, P(vp)]
        L.sgm_compute_disparity.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, vp, vp, C.c_ssize_t, i32, i32, vp, vp]
        L.sgm_compute_disparity_device.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, vp, vp, C.c_ssize_t, i32, i32, vp,
                                                   vp, vp]
        L.sgm_compute_disparity_raw.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, vp, vp, vp, vp]
        L.sgm_compute_disparity_raw_device.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, vp, vp, vp, vp, vp]
        L.sgm_point_params_default.argtypes = [P(CSgmPointParams)]
        L.sgm_disparity_to_points_device.argtypes = [vp, vp, i32, i32, P(CSgmRectifiedGeometry), P(CSgmPointParams),
                                                     P(CSgmFrame), vp]
        L.sgm_pack_depth_device.argtypes = [vp, vp, i32, i32, vp, vp]
        L.sgm_speckle_workspace_bytes.argtypes = [i32, i32]
        L.sgm_speckle_workspace_bytes.restype = C.c_size_t
        L.sgm_remove_speckles_device.argtypes = [vp, vp, i32, i32, i32, C.c_float, vp, vp]
        L.sgm_filter_disparity_device.argtypes = [vp, i32, i32, i32, C.c_float, i32, C.c_float, vp, vp]
        L.sgm_set_point_params.argtypes = [vp, P(CSgmRectifiedGeometry), P(CSgmPointParams)]
        L.sgm_compute_frame_raw.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, P(CSgmFrame), vp, vp]
        L.sgm_compute_frame_raw_device.argtypes = [vp, vp, vp, C.c_ssize_t, i32, i32, P(CSgmFrame), vp, vp, vp]
        L.sgm_get_stream.argtypes = [vp]
        L.sgm_get_stream.restype = vp
        L.sgm_synchronize.argtypes = [vp]
        L.sgm_last_timings.argtypes = [vp, P(CSgmTimings)]
        L.sgm_get_stage_totals.argtypes = [vp, P(CSgmStageTotals)]
        L.sgm_set_option.argtypes = [vp, i32, C.c_int64]
        L.sgm_reset_stage_totals.argtypes = [vp]
        L.sgm_device_bytes.argtypes = [vp]
        L.sgm_device_bytes.restype = C.c_size_t
        L.sgm_dev_malloc.argtypes = [C.c_size_t, P(vp)]
        L.sgm_dev_free.argtypes = [vp]
        L.sgm_dev_upload.argtypes = [vp, vp, C.c_size_t]
        L.sgm_dev_download.argtypes = [vp, vp, C.c_size_t]
        L.sgm_dev_memset.argtypes = [vp, i32, C.c_size_t]
        L.sgm_dev_copy.argtypes = [vp, vp, C.c_size_t]
        L.sgm_dev_mem_info.argtypes = [P(C.c_size_t), P(C.c_size_t)]
        L.sgm_host_alloc.argtypes = [C.c_size_t, P(vp)]
        L.sgm_host_free.argtypes = [vp]
        L.sgm_batch_create.argtypes = [P(CSgmParams), i32, P(CSgmPair), P(vp)]
        L.sgm_batch_create_tuned.argtypes = [P(CSgmParams), i32, P(CSgmPair), P(CSgmBatchTuning), P(vp)]
        L.sgm_batch_get_plan.argtypes = [vp, P(CSgmBatchPlan)]
        L.sgm_batch_run.argtypes = [vp, vp]
        L.sgm_batch_run_plain.argtypes = [vp, vp]
        L.sgm_batch_synchronize.argtypes = [vp]
        L.sgm_batch_last_error.argtypes = [vp]
        L.sgm_batch_last_error.restype = C.c_char_p
        L.sgm_batch_device_bytes.argtypes = [vp]
        L.sgm_batch_device_bytes.restype = C.c_size_t
        L.sgm_batch_destroy.argtypes = [vp]
        L.sgm_tune_cache_load.argtypes = [C.c_char_p, P(i32)]
        L.sgm_tune_cache_save.argtypes = [C.c_char_p, P(i32)]
        L.sgm_server_start.argtypes = [P(CSgmParams), P(CSgmServerConfig), P(vp)]
        L.sgm_server_submit.argtypes = [vp, P(CSgmRequest), P(C.c_uint64)]
        L.sgm_server_add_rectification.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, i32, P(C.c_int32)]
        L.sgm_server_poll.argtypes = [vp, P(CSgmCompletion), i32, i32, P(C.c_int)]
        L.sgm_server_get_stats.argtypes = [vp, P(CSgmServerStats)]
        L.sgm_server_last_error.argtypes = [vp]
        L.sgm_server_last_error.restype = C.c_char_p
        L.sgm_server_stop.argtypes = [vp]
        L.sgm_server_now_ns.argtypes = []
        L.sgm_server_now_ns.restype = C.c_int64
        _lib = L
    return _lib


def _check(rc: int, ctx=None, batch=None):
    if rc != 0:
        L = lib()
        if batch is not None:
            msg = L.sgm_batch_last_error(batch)
        else:
            msg = L.sgm_last_error(ctx)
        raise SgmError(rc, (msg or b"").decode(errors="replace"))


def _same_shape(a: np.ndarray, b: np.ndarray, what: str):
    """Both images 2-D and of one size: the C calls take one width, height and stride for both, so a
    smaller second image would be read past its end."""
    if a.ndim != 2 or a.shape != b.shape:
        raise SgmError(5, f"{what} must be 2-D images of the same size (got {a.shape} and {b.shape})")


def _ptr(a):
    return None if a is None else a.ctypes.data


class DeviceBuffer:
    """Device allocation owned by Python (sgm_dev_malloc / sgm_dev_free)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().sgm_dev_malloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        _check(lib().sgm_dev_upload(self.ptr + offset, _ptr(a), a.nbytes))

    def download(self, shape, dtype, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert offset + out.nbytes <= self.nbytes
        _check(lib().sgm_dev_download(_ptr(out), self.ptr + offset, out.nbytes))
        return out

    def memset(self, value: int = 0):
        _check(lib().sgm_dev_memset(self.ptr, value, self.nbytes))

    def free(self):
        if self.ptr:
            lib().sgm_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dev_copy(dst: int, src: int, nbytes: int):
    _check(lib().sgm_dev_copy(dst, src, nbytes))


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    _check(lib().sgm_dev_mem_info(C.byref(f), C.byref(t)))
    return f.value, t.value


class HipStereo:
    """gfx950 replacement of depth_metal::MetalStereo's SGM stage (metal_stereo.hpp:67-107)."""

    def __init__(self, handle, params: SgmParams, width: int, height: int):
        self._h = handle
        self.params = params
        self.width, self.height = width, height

    @classmethod
    def create(cls, params: SgmParams, width: int, height: int) -> "HipStereo":
        """MetalStereo::create (metal_stereo.hpp:70-71): validate and allocate for width x height."""
        h = C.c_void_p()
        cp = params.to_c()
        _check(lib().sgm_create(C.byref(cp), width, height, C.byref(h)))
        return cls(h.value, params, width, height)

    @classmethod
    def create_stereo(cls, params: SgmParams, stereo: StereoParams, width: int, height: int) -> "HipStereo":
        """MetalStereo::create with the full StereoParams: compute_disparity on width x height images,
        SGM at (width >> levels) x (height >> levels)."""
        h = C.c_void_p()
        cp = stereo.to_c(params)
        _check(lib().sgm_create_stereo(C.byref(cp), width, height, C.byref(h)))
        obj = cls(h.value, params, width, height)
        obj.stereo = stereo
        return obj

    def compute_disparity(self, left, right, full_left=None, full_right=None) -> dict:
        """depth::compute_disparity (stereo.cpp:496-536): disparity and confidence at left's size."""
        left, right = np.ascontiguousarray(left, np.uint8), np.ascontiguousarray(right, np.uint8)
        _same_shape(left, right, "left and right")
        h, w = left.shape
        fl = None if full_left is None else np.ascontiguousarray(full_left, np.uint8)
        fr = None if full_right is None else np.ascontiguousarray(full_right, np.uint8)
        if (fl is None) != (fr is None):
            raise SgmError(5, "full_left and full_right must be given together")
        if fl is not None:
            _same_shape(fl, fr, "full_left and full_right")
        fh, fw = (0, 0) if fl is None else fl.shape
        res = {"disparity": np.empty((h, w), np.float32), "confidence": np.empty((h, w), np.float32)}
        _check(lib().sgm_compute_disparity(self._h, _ptr(left), _ptr(right), w, w, h, _ptr(fl), _ptr(fr), fw, fw, fh,
                                           _ptr(res["disparity"]), _ptr(res["confidence"])), ctx=self._h)
        return res

    def compute_disparity_device(self, d_left: int, d_right: int, pitch: int, width: int, height: int,
                                 d_disparity: int, d_confidence: int | None = None, d_full_left: int | None = None,
                                 d_full_right: int | None = None, full_pitch: int = 0, stream: int | None = None):
        """Device-resident compute_disparity (full pair, when given, is 2*width x 2*height); no wait."""
        fw, fh = (2 * width, 2 * height) if d_full_left else (0, 0)
        _check(lib().sgm_compute_disparity_device(self._h, d_left, d_right, pitch, width, height, d_full_left,
                                                  d_full_right, full_pitch or fw, fw, fh, d_disparity, d_confidence,
                                                  stream), ctx=self._h)

    def compute_disparity_raw(self, raw_left, raw_right, rectified: bool = False) -> dict:
        """MetalStereo::compute_raw with the full StereoParams: rectify, then compute_disparity."""
        raw_left, raw_right = np.ascontiguousarray(raw_left, np.uint8), np.ascontiguousarray(raw_right, np.uint8)
        _same_shape(raw_left, raw_right, "raw frames")
        rh, rw = raw_left.shape
        w, h = getattr(self, "rect_size", (1, 1))
        res = {"disparity": np.empty((h, w), np.float32), "confidence": np.empty((h, w), np.float32)}
        rl = rr = None
        if rectified:
            res["rect_left"] = rl = np.empty((h, w), np.uint8)
            res["rect_right"] = rr = np.empty((h, w), np.uint8)
        _check(lib().sgm_compute_disparity_raw(self._h, _ptr(raw_left), _ptr(raw_right), rw, rw, rh,
                                               _ptr(res["disparity"]), _ptr(res["confidence"]), _ptr(rl), _ptr(rr)),
               ctx=self._h)
        return res

    def compute_disparity_raw_device(self, d_raw_left: int, d_raw_right: int, pitch: int, raw_width: int,
                                     raw_height: int, d_disparity: int, d_confidence: int | None = None,
                                     d_rect_left: int | None = None, d_rect_right: int | None = None,
                                     stream: int | None = None):
        _check(lib().sgm_compute_disparity_raw_device(self._h, d_raw_left, d_raw_right, pitch, raw_width, raw_height,
                                                      d_disparity, d_confidence, d_rect_left, d_rect_right, stream),
               ctx=self._h)

    def set_point_params(self, geometry: RectifiedGeometry, params: PointParams = PointParams()):
        """MetalStereo::set_point_params (metal_stereo.hpp:83-85)."""
        g, p = geometry.to_c(), params.to_c()
        _check(lib().sgm_set_point_params(self._h, C.byref(g), C.byref(p)), ctx=self._h)

    def compute_frame_raw(self, raw_left, raw_right, rectified: bool = False) -> dict:
        """MetalStereo::compute_frame_raw: raw frames -> points/normals (h, w, 4) and weights (h, w)."""
        raw_left, raw_right = np.ascontiguousarray(raw_left, np.uint8), np.ascontiguousarray(raw_right, np.uint8)
        _same_shape(raw_left, raw_right, "raw frames")
        rh, rw = raw_left.shape
        w, h = getattr(self, "rect_size", (1, 1))
        res = {"points": np.empty((h, w, 4), np.float32), "normals": np.empty((h, w, 4), np.float32),
               "weights": np.empty((h, w), np.float32)}
        rl = rr = None
        if rectified:
            res["rect_left"] = rl = np.empty((h, w), np.uint8)
            res["rect_right"] = rr = np.empty((h, w), np.uint8)
        fr = CSgmFrame(_ptr(res["points"]), _ptr(res["normals"]), _ptr(res["weights"]))
        _check(lib().sgm_compute_frame_raw(self._h, _ptr(raw_left), _ptr(raw_right), rw, rw, rh, C.byref(fr), _ptr(rl),
                                           _ptr(rr)), ctx=self._h)
        return res

    def compute_frame_raw_device(self, d_raw_left: int, d_raw_right: int, pitch: int, raw_width: int, raw_height: int,
                                 d_points: int | None = None, d_normals: int | None = None,
                                 d_weights: int | None = None, stream: int | None = None):
        fr = CSgmFrame(d_points, d_normals, d_weights)
        _check(lib().sgm_compute_frame_raw_device(self._h, d_raw_left, d_raw_right, pitch, raw_width, raw_height,
                                                  C.byref(fr), None, None, stream), ctx=self._h)

    def close(self):
        if self._h:
            lib().sgm_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _host_outputs(self, w: int, h: int, outputs) -> tuple[dict, CSgmOutputs]:
        res = {}
        co = CSgmOutputs()
        for f in outputs:
            shape = (h, w, self.params.num_disparities) if f in VOLUME_FIELDS else (h, w)
            res[f] = np.empty(shape, OUTPUT_DTYPES[f])
            setattr(co, f, res[f].ctypes.data)
        if "disparity" not in res:
            res["disparity"] = np.empty((h, w), np.float32)
            co.disparity = res["disparity"].ctypes.data
        return res, co

    def compute(self, left: np.ndarray, right: np.ndarray, outputs=("disparity",)) -> dict:
        """sgm_disparity on a rectified u8 pair (+ median3 etc. as requested in `outputs`)."""
        left = np.ascontiguousarray(left, dtype=np.uint8)
        right = np.ascontiguousarray(right, dtype=np.uint8)
        if left.shape != right.shape or left.ndim != 2:
            raise SgmError(5, "left and right must be 2-D images of the same size")
        h, w = left.shape
        res, co = self._host_outputs(w, h, outputs)
        _check(lib().sgm_compute(self._h, _ptr(left), _ptr(right), w, w, h, C.byref(co)), ctx=self._h)
        return res

    def set_rectification(self, left_map_x, left_map_y, right_map_x, right_map_y, raw_width: int, raw_height: int):
        """MetalStereo::set_rectification (metal_stereo.hpp:74-75): resident left/right remap tables
        (float32 map_x / map_y, each map_h x map_w); the stereo size becomes half the maps'."""
        maps = [np.ascontiguousarray(m, dtype=np.float32) for m in (left_map_x, left_map_y, right_map_x, right_map_y)]
        if any(m.ndim != 2 or m.shape != maps[0].shape for m in maps):
            raise SgmError(5, "the four maps must be 2-D arrays of one size")
        mh, mw = maps[0].shape
        _check(lib().sgm_set_rectification(self._h, *[_ptr(m) for m in maps], mw, mh, raw_width, raw_height), ctx=self._h)
        self.rect_size = (mw // 2, mh // 2)
        self.raw_size = (raw_width, raw_height)

    def compute_raw(self, raw_left: np.ndarray, raw_right: np.ndarray, outputs=("disparity",),
                    rectified: bool = False) -> dict:
        """MetalStereo::compute_raw (metal_stereo.hpp:79-81): rectify with the resident maps, then SGM.
        rectified=True also returns the rectified pair as rect_left / rect_right."""
        raw_left = np.ascontiguousarray(raw_left, dtype=np.uint8)
        raw_right = np.ascontiguousarray(raw_right, dtype=np.uint8)
        if raw_left.shape != raw_right.shape or raw_left.ndim != 2:
            raise SgmError(5, "raw frames must be 2-D images of the same size")
        rh, rw = raw_left.shape
        w, h = getattr(self, "rect_size", (0, 0))
        res, co = self._host_outputs(max(w, 1), max(h, 1), outputs)
        rl = rr = None
        if rectified:
            res["rect_left"] = rl = np.empty((max(h, 1), max(w, 1)), np.uint8)
            res["rect_right"] = rr = np.empty((max(h, 1), max(w, 1)), np.uint8)
        _check(lib().sgm_compute_raw(self._h, _ptr(raw_left), _ptr(raw_right), rw, rw, rh, C.byref(co), _ptr(rl),
                                     _ptr(rr)), ctx=self._h)
        return res

    def compute_raw_device(self, d_raw_left: int, d_raw_right: int, pitch: int, raw_width: int, raw_height: int,
                           d_outputs: dict | None = None, d_rect_left: int | None = None,
                           d_rect_right: int | None = None, stream: int | None = None):
        """Device-resident variant (sgm_compute_raw_device); returns without waiting."""
        co = CSgmOutputs()
        for f, p in (d_outputs or {}).items():
            setattr(co, f, p)
        _check(lib().sgm_compute_raw_device(self._h, d_raw_left, d_raw_right, pitch, raw_width, raw_height,
                                            C.byref(co), d_rect_left, d_rect_right, stream), ctx=self._h)

    def compute_device(self, d_left: int, d_right: int, pitch: int, width: int, height: int,
                       d_outputs: dict | None = None, stream: int | None = None):
        """Device-resident variant (sgm_compute_device); returns without waiting."""
        co = CSgmOutputs()
        for f, p in (d_outputs or {}).items():
            setattr(co, f, p)
        _check(lib().sgm_compute_device(self._h, d_left, d_right, pitch, width, height, C.byref(co), stream),
               ctx=self._h)

    def synchronize(self):
        _check(lib().sgm_synchronize(self._h), ctx=self._h)

    def stream(self) -> int:
        """The context's own stream handle (hipStream_t), to order other work after its calls."""
        return lib().sgm_get_stream(self._h)

    def last_timings(self) -> dict:
        """MetalStereo::last_timings (metal_stereo.hpp:97)."""
        t = CSgmTimings()
        _check(lib().sgm_last_timings(self._h, C.byref(t)), ctx=self._h)
        return {f: getattr(t, f) for f, _ in CSgmTimings._fields_}

    def stage_totals(self) -> dict:
        """Per-stage GPU ms summed over the calls completed since reset_stage_totals()."""
        t = CSgmStageTotals()
        _check(lib().sgm_get_stage_totals(self._h, C.byref(t)), ctx=self._h)
        return {f: getattr(t, f) for f, _ in CSgmStageTotals._fields_}

    def set_option(self, option: int, value: int):
        """Tuning knob (OPT_COST_VOLUME): changes speed and memory, never results."""
        _check(lib().sgm_set_option(self._h, option, value), ctx=self._h)

    def reset_stage_totals(self):
        _check(lib().sgm_reset_stage_totals(self._h), ctx=self._h)

    def device_bytes(self) -> int:
        return lib().sgm_device_bytes(self._h)


@dataclass
class PairSpec:
    """One pair of a batch: host images plus its disparity range."""

    left: np.ndarray
    right: np.ndarray
    min_disparity: int
    num_disparities: int
    outputs: tuple = ("disparity", "disparity_median", "disparity_int", "status", "right_disparity",
                      "best_cost", "second_cost")


class Batch:
    """A planned batch (sgm_batch_create): device copies of the pairs, outputs, and the graph."""

    def __init__(self, params: SgmParams, pairs: list[PairSpec], tuning: dict | None = None):
        """tuning: cost_volume / path_lines / branches / path_group / sweep_band / sweep_tile fixed to a
        value (omitted = autotuned)."""
        self.params = params
        self.specs = pairs
        self._bufs = []
        self._img_index = []  # position in _bufs of each pair's input buffer (left rows, then right)
        cps = (CSgmPair * len(pairs))()
        self.out_ptrs = []
        for i, sp in enumerate(pairs):
            h, w = sp.left.shape
            img = DeviceBuffer(2 * w * h)
            img.upload(np.ascontiguousarray(sp.left, np.uint8), 0)
            img.upload(np.ascontiguousarray(sp.right, np.uint8), w * h)
            self._img_index.append(len(self._bufs))
            self._bufs.append(img)
            cps[i].left = img.ptr
            cps[i].right = img.ptr + w * h
            cps[i].pitch = w
            cps[i].width = w
            cps[i].height = h
            cps[i].min_disparity = sp.min_disparity
            cps[i].num_disparities = sp.num_disparities
            outs = {}
            for f in set(sp.outputs) | {"disparity"}:
                shape = (h, w, sp.num_disparities) if f in VOLUME_FIELDS else (h, w)
                b = DeviceBuffer(int(np.prod(shape)) * np.dtype(OUTPUT_DTYPES[f]).itemsize)
                self._bufs.append(b)
                setattr(cps[i].out, f, b.ptr)
                outs[f] = (b, shape, OUTPUT_DTYPES[f])
            self.out_ptrs.append(outs)
        h = C.c_void_p()
        cp = params.to_c()
        t = tuning or {}
        tn = CSgmBatchTuning(t.get("cost_volume", -1), t.get("path_lines", -1), t.get("branches", -1),
                             t.get("path_group", -1), t.get("sweep_band", -1), t.get("sweep_tile", -1))
        rc = lib().sgm_batch_create_tuned(C.byref(cp), len(pairs), cps, C.byref(tn), C.byref(h))
        if rc != 0:
            raise SgmError(rc, (lib().sgm_batch_last_error(None) or b"").decode())
        self._h = h.value

    def plan(self) -> dict:
        """The launch plan the autotuner chose (sgm_batch_get_plan)."""
        pl = CSgmBatchPlan()
        _check(lib().sgm_batch_get_plan(self._h, C.byref(pl)), batch=self._h)
        return {f: getattr(pl, f) for f, _ in CSgmBatchPlan._fields_}

    def run(self, stream=None):
        _check(lib().sgm_batch_run(self._h, stream), batch=self._h)

    def run_plain(self, stream=None):
        _check(lib().sgm_batch_run_plain(self._h, stream), batch=self._h)

    def synchronize(self):
        _check(lib().sgm_batch_synchronize(self._h), batch=self._h)

    def set_inputs(self, i: int, left: np.ndarray, right: np.ndarray):
        """Replaces pair i's images (same size) in device memory; the next run uses them."""
        h, w = self.specs[i].left.shape
        assert left.shape == (h, w) and right.shape == (h, w)
        buf = self._bufs[self._img_index[i]]
        buf.upload(np.ascontiguousarray(left, np.uint8), 0)
        buf.upload(np.ascontiguousarray(right, np.uint8), w * h)

    def results(self, i: int) -> dict:
        return {f: b.download(shape, dt) for f, (b, shape, dt) in self.out_ptrs[i].items()}

    def clear_outputs(self, value: int = 0):
        for outs in self.out_ptrs:
            for b, _, _ in outs.values():
                b.memset(value)

    def device_bytes(self) -> int:
        return lib().sgm_batch_device_bytes(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().sgm_batch_destroy(self._h)
            self._h = None
        for b in self._bufs:
            b.free()
        self._bufs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_cache_load(path: str) -> int:
    """Merges a plan-cache file (tune_cache_save) into the process's autotuner cache; returns its entries."""
    n = C.c_int(0)
    rc = lib().sgm_tune_cache_load(os.fsencode(path), C.byref(n))
    if rc != 0:
        raise SgmError(rc, (lib().sgm_last_error(None) or b"").decode())
    return n.value


def tune_cache_save(path: str) -> int:
    """Writes the process's autotuner plan cache to `path`; returns the number of entries."""
    n = C.c_int(0)
    rc = lib().sgm_tune_cache_save(os.fsencode(path), C.byref(n))
    if rc != 0:
        raise SgmError(rc, (lib().sgm_last_error(None) or b"").decode())
    return n.value


class PinnedPool:
    """Reusable pinned host arrays (sgm_host_alloc), so device copies into them are direct DMA.
    Thread-safe."""

    def __init__(self):
        self._free = {}   # (nbytes) -> [ptr]
        self._all = []
        self._lock = threading.Lock()

    def take(self, shape, dtype) -> np.ndarray:
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        with self._lock:
            lst = self._free.get(nbytes)
            ptr = lst.pop() if lst else None
        if ptr is None:
            p = C.c_void_p()
            _check(lib().sgm_host_alloc(nbytes, C.byref(p)))
            ptr = p.value
            with self._lock:
                self._all.append(ptr)
        buf = (C.c_uint8 * nbytes).from_address(ptr)
        a = np.frombuffer(buf, dtype=dtype).reshape(shape)
        return a

    def give(self, a: np.ndarray):
        with self._lock:
            self._free.setdefault(a.nbytes, []).append(a.ctypes.data)

    def free_all(self):
        with self._lock:
            ptrs, self._all, self._free = self._all, [], {}
        for p in ptrs:
            lib().sgm_host_free(p)


class Server:
    """The serving loop (sgm_server_*): submit host pairs as they arrive, poll completions.

    Output arrays live in pinned host memory owned by this object; a completion's arrays stay valid
    until `release(completion)` returns them to the pool (or the server stops). submit, poll and
    release may be called from any threads (the library's server is thread-safe; a lock here keeps a
    ticket's output arrays registered before any poll can return its completion)."""

    def __init__(self, params: SgmParams, max_width: int, max_height: int, max_disparities: int, slots: int = 4,
                 hbm_budget_bytes: int = 0, compute_streams: int = 0, path_group: int = 0):
        """compute_streams: 0 = one shared compute stream (requests' kernels in admission order, copies
        overlapping them), k = k shared streams, -1 = each slot computes on its own stream.
        path_group: lanes per path scanline, 0 = timed per request shape, or 16 / 32 / 64."""
        cfg = CSgmServerConfig(max_width, max_height, max_disparities, slots, hbm_budget_bytes, compute_streams,
                               path_group)
        h = C.c_void_p()
        cp = params.to_c()
        rc = lib().sgm_server_start(C.byref(cp), C.byref(cfg), C.byref(h))
        if rc != 0:
            raise SgmError(rc, (lib().sgm_last_error(None) or b"").decode())
        self._h = h.value
        self._pending = {}  # ticket -> outputs dict
        self._pool = PinnedPool()
        self._rect_sizes = {}  # map set id -> rectified (width, height)
        self._borrowed = {}    # ticket -> input arrays read in place, kept alive until polled
        self._lock = threading.Lock()  # guards _pending and _borrowed across submit and poll

    @staticmethod
    def now_ns() -> int:
        return lib().sgm_server_now_ns()

    def add_rectification(self, left_map_x, left_map_y, right_map_x, right_map_y, raw_width: int,
                          raw_height: int) -> int:
        """Registers a resident map set (sgm_server_add_rectification); returns its id for submit()."""
        maps = [np.ascontiguousarray(m, dtype=np.float32) for m in (left_map_x, left_map_y, right_map_x, right_map_y)]
        if any(m.ndim != 2 or m.shape != maps[0].shape for m in maps):
            raise SgmError(5, "the four maps must be 2-D arrays of one size")
        mh, mw = maps[0].shape
        i = C.c_int32()
        rc = lib().sgm_server_add_rectification(self._h, *[m.ctypes.data for m in maps], mw, mh, raw_width, raw_height,
                                                C.byref(i))
        if rc != 0:
            raise SgmError(rc, (lib().sgm_server_last_error(self._h) or b"").decode())
        self._rect_sizes[i.value] = (mw // 2, mh // 2)
        return i.value

    def submit(self, left: np.ndarray, right: np.ndarray, min_disparity: int, num_disparities: int,
               outputs=("disparity",), tag: int = 0, rectification: int = 0, borrow: bool = False) -> int:
        """A rectified pair, or (rectification = a map set id) a raw pair rectified on the device.
        borrow=True: the arrays are read in place and must stay unchanged until the completion is
        polled (contiguous uint8 arrays, ideally in pinned memory)."""
        if borrow and not all(a.dtype == np.uint8 and a.flags.c_contiguous for a in (left, right)):
            raise SgmError(5, "borrowed inputs must be C-contiguous uint8 arrays")
        left = np.ascontiguousarray(left, dtype=np.uint8)
        right = np.ascontiguousarray(right, dtype=np.uint8)
        ih, iw = left.shape
        w, h = self._rect_sizes.get(rectification, (iw, ih)) if rectification else (iw, ih)
        res = {f: self._pool.take((h, w), OUTPUT_DTYPES[f]) for f in set(outputs) | {"disparity"}}
        rq = CSgmRequest()
        rq.left, rq.right, rq.stride = left.ctypes.data, right.ctypes.data, iw
        rq.width, rq.height, rq.min_disparity, rq.num_disparities = w, h, min_disparity, num_disparities
        for f, a in res.items():
            setattr(rq.out, f, a.ctypes.data)
        rq.user_tag = tag
        rq.rectification = rectification
        rq.borrow_inputs = int(borrow)
        t = C.c_uint64()
        # The completion can be polled as soon as the request is submitted: its outputs are
        # registered under the lock that poll takes before it looks them up.
        with self._lock:
            rc = lib().sgm_server_submit(self._h, C.byref(rq), C.byref(t))
            if rc == 0:
                self._pending[t.value] = res
                if borrow:
                    self._borrowed[t.value] = (left, right)
        if rc != 0:
            for a in res.values():
                self._pool.give(a)
            raise SgmError(rc, (lib().sgm_server_last_error(self._h) or b"").decode())
        return t.value

    def poll(self, max_n: int = 64, timeout_us: int = 0) -> list:
        """Completions as dicts (ticket, tag, status, t_submit_ns, t_admit_ns, t_done_ns, gpu_ms, outputs)."""
        arr = (CSgmCompletion * max_n)()
        n = C.c_int()
        _check(lib().sgm_server_poll(self._h, arr, max_n, timeout_us, C.byref(n)))
        out = []
        with self._lock:
            for i in range(n.value):
                c = arr[i]
                out.append({"ticket": c.ticket, "tag": c.user_tag, "status": c.status, "t_submit_ns": c.t_submit_ns,
                            "t_admit_ns": c.t_admit_ns, "t_done_ns": c.t_done_ns, "gpu_ms": c.gpu_ms,
                            "outputs": self._pending.pop(c.ticket, None)})
                self._borrowed.pop(c.ticket, None)
        return out

    def release(self, completion: dict):
        """Returns a polled completion's output arrays to the pinned pool (they must not be used after)."""
        for a in (completion.get("outputs") or {}).values():
            self._pool.give(a)
        completion["outputs"] = None

    def stats(self) -> dict:
        st = CSgmServerStats()
        _check(lib().sgm_server_get_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in CSgmServerStats._fields_}

    def stop(self):
        if getattr(self, "_h", None):
            lib().sgm_server_stop(self._h)
            self._h = None
            self._pool.free_all()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def disparity_to_points(disparity, confidence, geometry: RectifiedGeometry, params: PointParams = PointParams()) -> dict:
    """depth::disparity_to_points on the GPU (sgm_disparity_to_points_device) for host arrays:
    points and normals (h, w, 4) as (x, y, z, valid), weights (h, w)."""
    d = np.ascontiguousarray(disparity, np.float32)
    h, w = d.shape
    bufs = [DeviceBuffer(d.nbytes), DeviceBuffer(d.nbytes) if confidence is not None else None,
            DeviceBuffer(16 * w * h), DeviceBuffer(16 * w * h), DeviceBuffer(4 * w * h)]
    try:
        bufs[0].upload(d, 0)
        if confidence is not None:
            bufs[1].upload(np.ascontiguousarray(confidence, np.float32), 0)
        fr = CSgmFrame(bufs[2].ptr, bufs[3].ptr, bufs[4].ptr)
        g, p = geometry.to_c(), params.to_c()
        rc = lib().sgm_disparity_to_points_device(bufs[0].ptr, bufs[1].ptr if bufs[1] else None, w, h, C.byref(g),
                                                  C.byref(p), C.byref(fr), None)
        if rc != 0:
            raise SgmError(rc, (lib().sgm_last_error(None) or b"").decode())
        return {"points": bufs[2].download((h, w, 4), np.float32), "normals": bufs[3].download((h, w, 4), np.float32),
                "weights": bufs[4].download((h, w), np.float32)}
    finally:
        for b in bufs:
            if b is not None:
                b.free()


def remove_speckles(disparity, max_region_size: int = 200, max_diff: float = 1.0) -> np.ndarray:
    """depth::remove_speckles on the GPU (sgm_remove_speckles_device) for a host array: a copy with every
    region (4-neighbourhood, |d_a - d_b| <= max_diff) of fewer than max_region_size pixels invalid."""
    d = np.ascontiguousarray(disparity, np.float32)
    h, w = d.shape
    bd = DeviceBuffer(d.nbytes)
    bw = DeviceBuffer(max(4, lib().sgm_speckle_workspace_bytes(w, h)))
    try:
        bd.upload(d, 0)
        _check(lib().sgm_remove_speckles_device(bd.ptr, bd.ptr, w, h, int(max_region_size), float(max_diff), bw.ptr, None))
        return bd.download((h, w), np.float32)
    finally:
        bd.free()
        bw.free()


def filter_disparity(disparity, occlusion: bool = True, occlusion_tolerance: float = 1.0, spikes: bool = True,
                     spike_threshold: float = 1.0) -> np.ndarray:
    """depth::filter_disparity on the GPU (sgm_filter_disparity_device) for a host array: a filtered copy
    (occlusion filter, then spike filter, with the reference's FilterParams)."""
    d = np.ascontiguousarray(disparity, np.float32)
    h, w = d.shape
    bd, bw = DeviceBuffer(d.nbytes), DeviceBuffer(d.nbytes)
    try:
        bd.upload(d, 0)
        _check(lib().sgm_filter_disparity_device(bd.ptr, w, h, int(bool(occlusion)), float(occlusion_tolerance),
                                                 int(bool(spikes)), float(spike_threshold), bw.ptr, None))
        return bd.download((h, w), np.float32)
    finally:
        bd.free()
        bw.free()


def pack_depth_device(d_points: int, d_weights: int | None, width: int, height: int, d_out: int,
                      stream: int | None = None):
    """DepthPacker::pack on device data (sgm_pack_depth_device), enqueued on `stream`; no wait."""
    rc = lib().sgm_pack_depth_device(d_points, d_weights, width, height, d_out, stream)
    if rc != 0:
        raise SgmError(rc, (lib().sgm_last_error(None) or b"").decode())


def pack_depth(points_xyzw, weights=None) -> np.ndarray:
    """DepthPacker::pack on the GPU (sgm_pack_depth_device) for host arrays: 3*w*h bytes."""
    pts = np.ascontiguousarray(points_xyzw, np.float32)
    h, w = pts.shape[:2]
    bp, bo = DeviceBuffer(pts.nbytes), DeviceBuffer(3 * w * h)
    bw = DeviceBuffer(4 * w * h) if weights is not None else None
    try:
        bp.upload(pts, 0)
        if bw is not None:
            bw.upload(np.ascontiguousarray(weights, np.float32), 0)
        rc = lib().sgm_pack_depth_device(bp.ptr, bw.ptr if bw else None, w, h, bo.ptr, None)
        if rc != 0:
            raise SgmError(rc, (lib().sgm_last_error(None) or b"").decode())
        return bo.download((3 * w * h,), np.uint8)
    finally:
        for b in (bp, bo, bw):
            if b is not None:
                b.free()
