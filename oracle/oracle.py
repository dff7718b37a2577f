"""ctypes bindings of the CPU oracle (``liboracle_sgm.so``) -- test infrastructure only.

The oracle restates the reference CPU SGM (libs/depth/src/stereo.cpp:15-96, 129-147, 328-346,
361-463) in C. Only the tests, the smoke check and the CPU-baseline leg of the benchmark use it,
as the checker; the library never does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sgm.so")

INVALID_DISPARITY = np.float32(-1.0e9)
NO_MATCH = np.int32(-(2**31))
ST_WINDOW, ST_UNIQUE, ST_LR, ST_VALID = 1, 2, 4, 7


class OrcParams(C.Structure):
    _fields_ = [
        ("min_disparity", C.c_int32),
        ("num_disparities", C.c_int32),
        ("census_radius_x", C.c_int32),
        ("census_radius_y", C.c_int32),
        ("p1", C.c_int32),
        ("p2", C.c_int32),
        ("adaptive_p2", C.c_int32),
        ("eight_paths", C.c_int32),
        ("slant_steps", C.c_int32),
        ("uniqueness", C.c_float),
        ("lr_max_diff", C.c_int32),
        ("subpixel", C.c_int32),
    ]


class OrcOutputs(C.Structure):
    _fields_ = [
        ("disparity", C.c_void_p),
        ("disparity_median", C.c_void_p),
        ("disparity_int", C.c_void_p),
        ("status", C.c_void_p),
        ("right_disparity", C.c_void_p),
        ("best_cost", C.c_void_p),
        ("second_cost", C.c_void_p),
        ("sum_volume", C.c_void_p),
        ("cost_volume", C.c_void_p),
        ("census_left", C.c_void_p),
        ("census_right", C.c_void_p),
    ]


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

    def to_c(self) -> OrcParams:
        return OrcParams(self.min_disparity, self.num_disparities, self.census_radius_x, self.census_radius_y,
                         self.p1, self.p2, int(self.adaptive_p2), int(self.eight_paths), int(self.slant_steps),
                         float(self.uniqueness), self.lr_max_diff, int(self.subpixel))


class OrcStereoParams(C.Structure):
    _fields_ = [
        ("pyramid_levels", C.c_int32),
        ("refine_search_radius", C.c_int32), ("refine_zncc_radius", C.c_int32), ("refine_min_zncc", C.c_float),
        ("refine_reject_edge_winners", C.c_int32),
        ("detail_enabled", C.c_int32), ("detail_radius", C.c_int32), ("detail_search", C.c_float),
        ("detail_step", C.c_float), ("detail_min_zncc", C.c_float),
        ("filter_occlusion", C.c_int32), ("filter_occlusion_tolerance", C.c_float), ("filter_spikes", C.c_int32),
        ("filter_spike_threshold", C.c_float),
        ("speckle_max_region_size", C.c_int32), ("speckle_max_diff", C.c_float),
    ]


@dataclass
class StereoParams:
    """depth::StereoParams (libs/depth/include/einstar/depth/stereo.hpp:42-89) minus sgm/subpixel,
    with the reference defaults."""

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

    def to_c(self) -> OrcStereoParams:
        return OrcStereoParams(*[type_(getattr(self, f)) if type_ is float else int(getattr(self, f))
                                 for f, type_ in ((f, float if t is C.c_float else int)
                                                  for f, t in OrcStereoParams._fields_)])


class OrcGeometry(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("f", "cx", "cy", "baseline", "cx_offset")]


class OrcPointParams(C.Structure):
    _fields_ = [(f, C.c_float) for f in ("min_depth", "max_depth", "max_depth_jump", "edge_weight",
                                         "weight_reference_depth")]


@dataclass
class Geometry:
    """depth::RectifiedGeometry (libs/depth/include/einstar/depth/point_image.hpp:12-26)."""

    f: float
    cx: float
    cy: float
    baseline: float
    cx_offset: float = 0.0


@dataclass
class PointParams:
    """depth::PointImageParams (point_image.hpp:43-49) with the reference defaults."""

    min_depth: float = 150.0
    max_depth: float = 700.0
    max_depth_jump: float = 4.0
    edge_weight: float = 0.25
    weight_reference_depth: float = 0.0


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        vp, i32, f32 = C.c_void_p, C.c_int, C.c_float
        L.orc_census.argtypes = [vp, i32, i32, C.c_ssize_t, i32, i32, vp, i32]
        L.orc_cost_volume.argtypes = [vp, vp, i32, i32, i32, i32, C.c_uint16, vp, i32]
        L.orc_aggregate_direction.argtypes = [vp, vp, C.c_ssize_t, i32, i32, i32, C.POINTER(OrcParams), i32, i32, vp, i32]
        L.orc_aggregate.argtypes = [vp, vp, C.c_ssize_t, i32, i32, C.POINTER(OrcParams), vp, i32]
        L.orc_subpixel_offset.argtypes = [f32, f32, f32]
        L.orc_subpixel_offset.restype = f32
        L.orc_wta.argtypes = [vp, i32, i32, C.POINTER(OrcParams), vp, vp, vp, vp, vp, vp, i32]
        L.orc_lr_check.argtypes = [vp, vp, i32, i32, i32, vp, i32]
        L.orc_median3.argtypes = [vp, i32, i32, vp, i32]
        L.orc_sgm.argtypes = [vp, vp, C.c_ssize_t, i32, i32, C.POINTER(OrcParams), C.POINTER(OrcOutputs), i32]
        L.orc_remap.argtypes = [vp, i32, i32, C.c_ssize_t, vp, vp, i32, i32, vp, i32]
        L.orc_downsample2.argtypes = [vp, i32, i32, C.c_ssize_t, vp, i32]
        L.orc_rectify_half.argtypes = [vp, i32, i32, C.c_ssize_t, vp, vp, i32, i32, vp, i32]
        L.orc_rectify_half.restype = i32
        L.orc_sgm.restype = i32
        SP = C.POINTER(OrcStereoParams)
        L.orc_zncc_slanted.argtypes = [vp, vp, C.c_ssize_t, i32, i32, i32, f32, f32, f32, i32]
        L.orc_zncc_slanted.restype = f32
        L.orc_zncc_window.argtypes = [vp, vp, C.c_ssize_t, i32, i32, i32, i32, f32, f32, f32, f32, f32]
        L.orc_zncc_window.restype = f32
        L.orc_refine_level.argtypes = [vp, vp, C.c_ssize_t, i32, i32, vp, i32, i32, SP, i32, vp, vp, i32]
        L.orc_refine_detail.argtypes = [vp, vp, C.c_ssize_t, i32, i32, vp, i32, i32, SP, i32, i32]
        L.orc_filter_disparity.argtypes = [vp, i32, i32, SP, i32]
        L.orc_remove_speckles.argtypes = [vp, i32, i32, SP]
        L.orc_remove_speckles.restype = i32
        L.orc_compute_disparity.argtypes = [vp, vp, C.c_ssize_t, i32, i32, vp, vp, C.c_ssize_t, i32, i32,
                                            C.POINTER(OrcParams), SP, vp, vp, i32]
        L.orc_compute_disparity.restype = i32
        L.orc_disparity_to_points.argtypes = [vp, vp, i32, i32, C.POINTER(OrcGeometry), C.POINTER(OrcPointParams), vp, vp,
                                              vp, i32]
        L.orc_pack_depth.argtypes = [vp, vp, i32, i32, vp]
        L.orc_unpack_depth.argtypes = [vp, i32, i32, vp, vp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def census(img: np.ndarray, rx: int, ry: int, threads: int = 0) -> np.ndarray:
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape
    out = np.empty((h, w), np.uint64)
    lib().orc_census(_ptr(img), w, h, w, rx, ry, _ptr(out), threads)
    return out


def cost_volume(cl, cr, min_d: int, nd: int, invalid_cost: int, threads: int = 0) -> np.ndarray:
    cl = np.ascontiguousarray(cl, dtype=np.uint64)
    cr = np.ascontiguousarray(cr, dtype=np.uint64)
    h, w = cl.shape
    out = np.empty((h, w, nd), np.uint16)
    lib().orc_cost_volume(_ptr(cl), _ptr(cr), w, h, min_d, nd, invalid_cost, _ptr(out), threads)
    return out


def aggregate_direction(cost, guide, p: SgmParams, dx: int, dy: int, sum_in=None, threads: int = 0) -> np.ndarray:
    cost = np.ascontiguousarray(cost, dtype=np.uint16)
    guide = np.ascontiguousarray(guide, dtype=np.uint8)
    h, w, nd = cost.shape
    s = np.zeros_like(cost) if sum_in is None else np.array(sum_in, dtype=np.uint16, copy=True)
    cp = p.to_c()
    lib().orc_aggregate_direction(_ptr(cost), _ptr(guide), w, w, h, nd, C.byref(cp), dx, dy, _ptr(s), threads)
    return s


def aggregate(cost, guide, p: SgmParams, threads: int = 0) -> np.ndarray:
    cost = np.ascontiguousarray(cost, dtype=np.uint16)
    guide = np.ascontiguousarray(guide, dtype=np.uint8)
    h, w, nd = cost.shape
    assert nd == p.num_disparities
    s = np.zeros_like(cost)
    cp = p.to_c()
    lib().orc_aggregate(_ptr(cost), _ptr(guide), w, w, h, C.byref(cp), _ptr(s), threads)
    return s


def subpixel_offset(cm: float, c0: float, cp: float) -> float:
    return lib().orc_subpixel_offset(cm, c0, cp)


def wta(sum_volume, p: SgmParams, threads: int = 0) -> dict:
    s = np.ascontiguousarray(sum_volume, dtype=np.uint16)
    h, w, nd = s.shape
    assert nd == p.num_disparities
    o = {
        "disparity": np.empty((h, w), np.float32),
        "right_disparity": np.empty((h, w), np.int32),
        "disparity_int": np.empty((h, w), np.int32),
        "status": np.empty((h, w), np.uint8),
        "best_cost": np.empty((h, w), np.uint16),
        "second_cost": np.empty((h, w), np.uint16),
    }
    cp = p.to_c()
    lib().orc_wta(_ptr(s), w, h, C.byref(cp), _ptr(o["disparity"]), _ptr(o["right_disparity"]),
                  _ptr(o["disparity_int"]), _ptr(o["status"]), _ptr(o["best_cost"]), _ptr(o["second_cost"]), threads)
    return o


def lr_check(disp, right_disp, lr_max_diff: int, status=None, threads: int = 0):
    d = np.array(disp, dtype=np.float32, copy=True)
    r = np.ascontiguousarray(right_disp, dtype=np.int32)
    st = None if status is None else np.array(status, dtype=np.uint8, copy=True)
    h, w = d.shape
    lib().orc_lr_check(_ptr(d), _ptr(r), w, h, lr_max_diff, _ptr(st), threads)
    return d if st is None else (d, st)


def median3(disp, threads: int = 0) -> np.ndarray:
    d = np.ascontiguousarray(disp, dtype=np.float32)
    h, w = d.shape
    out = np.empty_like(d)
    lib().orc_median3(_ptr(d), w, h, _ptr(out), threads)
    return out


def sgm(left, right, p: SgmParams, volumes: bool = False, threads: int = 0) -> dict:
    """sgm_disparity + median3 with every intermediate output; volumes=True also returns C and S."""
    left = np.ascontiguousarray(left, dtype=np.uint8)
    right = np.ascontiguousarray(right, dtype=np.uint8)
    assert left.shape == right.shape
    h, w = left.shape
    nd = p.num_disparities
    o = {
        "disparity": np.empty((h, w), np.float32),
        "disparity_median": np.empty((h, w), np.float32),
        "disparity_int": np.empty((h, w), np.int32),
        "status": np.empty((h, w), np.uint8),
        "right_disparity": np.empty((h, w), np.int32),
        "best_cost": np.empty((h, w), np.uint16),
        "second_cost": np.empty((h, w), np.uint16),
        "census_left": np.empty((h, w), np.uint64),
        "census_right": np.empty((h, w), np.uint64),
    }
    if volumes:
        o["sum_volume"] = np.empty((h, w, nd), np.uint16)
        o["cost_volume"] = np.empty((h, w, nd), np.uint16)
    co = OrcOutputs(*[_ptr(o.get(f[0])) for f in OrcOutputs._fields_])
    cp = p.to_c()
    rc = lib().orc_sgm(_ptr(left), _ptr(right), w, w, h, C.byref(cp), C.byref(co), threads)
    if rc != 0:
        raise MemoryError("oracle sgm: allocation failed")
    return o


def remap(src, map_x, map_y, threads: int = 0) -> np.ndarray:
    """calib::remap (libs/calib/src/rectify.cpp:133-148)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    mx = np.ascontiguousarray(map_x, dtype=np.float32)
    my = np.ascontiguousarray(map_y, dtype=np.float32)
    sh, sw = src.shape
    h, w = mx.shape
    out = np.empty((h, w), np.uint8)
    lib().orc_remap(_ptr(src), sw, sh, sw, _ptr(mx), _ptr(my), w, h, _ptr(out), threads)
    return out


def downsample2(src, threads: int = 0) -> np.ndarray:
    """downsample2 (libs/depth/src/stereo.cpp:348-359)."""
    src = np.ascontiguousarray(src, dtype=np.uint8)
    h, w = src.shape
    out = np.empty((h // 2, w // 2), np.uint8)
    lib().orc_downsample2(_ptr(src), w, h, w, _ptr(out), threads)
    return out


def rectify_half(raw, map_x, map_y, threads: int = 0) -> np.ndarray:
    """remap to the maps' resolution, then downsample2 (the SGM-level rectified image)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    mx = np.ascontiguousarray(map_x, dtype=np.float32)
    my = np.ascontiguousarray(map_y, dtype=np.float32)
    rh, rw = raw.shape
    h, w = mx.shape
    out = np.empty((h // 2, w // 2), np.uint8)
    if lib().orc_rectify_half(_ptr(raw), rw, rh, rw, _ptr(mx), _ptr(my), w, h, _ptr(out), threads) != 0:
        raise MemoryError("oracle rectify_half: allocation failed")
    return out


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def zncc_slanted(a, b, x, y, d, gx, gy, r) -> float:
    """zncc_slanted (libs/depth/src/stereo.cpp:101-125)."""
    a, b = _u8(a), _u8(b)
    assert a.shape == b.shape
    return lib().orc_zncc_slanted(_ptr(a), _ptr(b), a.shape[1], a.shape[1], x, y, d, gx, gy, r)


def zncc_window(a, b, x0, y0, n, cx, cy, D, gx, gy) -> float:
    """zncc_window (libs/depth/src/stereo.cpp:211-239)."""
    a, b = _u8(a), _u8(b)
    assert a.shape == b.shape
    return lib().orc_zncc_window(_ptr(a), _ptr(b), a.shape[1], a.shape[1], x0, y0, n, cx, cy, D, gx, gy)


def refine_level(left, right, coarse_raw, sp: StereoParams, subpixel=True, threads: int = 0):
    """refine_level (libs/depth/src/stereo.cpp:151-207): (disparity, confidence) at left's size."""
    left, right = _u8(left), _u8(right)
    coarse = np.ascontiguousarray(coarse_raw, dtype=np.float32)
    h, w = left.shape
    ch, cw = coarse.shape
    d = np.empty((h, w), np.float32)
    c = np.empty((h, w), np.float32)
    csp = sp.to_c()
    lib().orc_refine_level(_ptr(left), _ptr(right), w, w, h, _ptr(coarse), cw, ch, C.byref(csp), int(subpixel),
                           _ptr(d), _ptr(c), threads)
    return d, c


def refine_detail(full_left, full_right, disparity, sp: StereoParams, subpixel=True, threads: int = 0):
    """refine_detail (libs/depth/src/stereo.cpp:243-284) on a copy of disparity."""
    fl, fr = _u8(full_left), _u8(full_right)
    d = np.array(disparity, dtype=np.float32, copy=True)
    H, W = fl.shape
    h, w = d.shape
    csp = sp.to_c()
    lib().orc_refine_detail(_ptr(fl), _ptr(fr), W, W, H, _ptr(d), w, h, C.byref(csp), int(subpixel), threads)
    return d


def filter_disparity(disparity, sp: StereoParams, threads: int = 0):
    """filter_disparity (libs/depth/src/stereo.cpp:286-326) on a copy."""
    d = np.array(disparity, dtype=np.float32, copy=True)
    h, w = d.shape
    csp = sp.to_c()
    lib().orc_filter_disparity(_ptr(d), w, h, C.byref(csp), threads)
    return d


def remove_speckles(disparity, sp: StereoParams):
    """remove_speckles (libs/depth/src/stereo.cpp:465-494) on a copy."""
    d = np.array(disparity, dtype=np.float32, copy=True)
    h, w = d.shape
    csp = sp.to_c()
    if lib().orc_remove_speckles(_ptr(d), w, h, C.byref(csp)) != 0:
        raise MemoryError("oracle remove_speckles: allocation failed")
    return d


def compute_disparity(left, right, sgm_params: SgmParams, sp: StereoParams, full_left=None, full_right=None,
                      threads: int = 0) -> dict:
    """compute_disparity (libs/depth/src/stereo.cpp:496-536), both forms."""
    left, right = _u8(left), _u8(right)
    h, w = left.shape
    fl = None if full_left is None else _u8(full_left)
    fr = None if full_right is None else _u8(full_right)
    fh, fw = (0, 0) if fl is None else fl.shape
    disp = np.empty((h, w), np.float32)
    conf = np.empty((h, w), np.float32)
    cp, csp = sgm_params.to_c(), sp.to_c()
    rc = lib().orc_compute_disparity(_ptr(left), _ptr(right), w, w, h, _ptr(fl), _ptr(fr), fw, fw, fh, C.byref(cp),
                                     C.byref(csp), _ptr(disp), _ptr(conf), threads)
    if rc != 0:
        raise MemoryError("oracle compute_disparity failed")
    return {"disparity": disp, "confidence": conf}


def disparity_to_points(disparity, confidence, g: Geometry, p: PointParams = PointParams(), threads: int = 0) -> dict:
    """disparity_to_points (libs/depth/src/point_image.cpp:10-55): points, normals [h, w, 3], weights."""
    d = np.ascontiguousarray(disparity, dtype=np.float32)
    c = None if confidence is None else np.ascontiguousarray(confidence, dtype=np.float32)
    h, w = d.shape
    out = {"points": np.empty((h, w, 3), np.float32), "normals": np.empty((h, w, 3), np.float32),
           "weights": np.empty((h, w), np.float32)}
    cg = OrcGeometry(g.f, g.cx, g.cy, g.baseline, g.cx_offset)
    cp = OrcPointParams(p.min_depth, p.max_depth, p.max_depth_jump, p.edge_weight, p.weight_reference_depth)
    lib().orc_disparity_to_points(_ptr(d), _ptr(c), w, h, C.byref(cg), C.byref(cp), _ptr(out["points"]),
                                  _ptr(out["normals"]), _ptr(out["weights"]), threads)
    return out


def pack_depth(points, weights) -> np.ndarray:
    """The session file's uncompressed depth layout (libs/session/src/session.cpp:143-160): 3*w*h bytes."""
    pts = np.ascontiguousarray(points, dtype=np.float32)
    wt = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    h, w = pts.shape[:2]
    out = np.empty(3 * w * h, np.uint8)
    lib().orc_pack_depth(_ptr(pts), _ptr(wt), w, h, _ptr(out))
    return out


def unpack_depth(packed, w: int, h: int):
    """decode_images without zstd (session.cpp:164-185): (depth mm, confidence)."""
    b = np.ascontiguousarray(packed, dtype=np.uint8)
    depth = np.empty((h, w), np.float32)
    conf = np.empty((h, w), np.float32)
    lib().orc_unpack_depth(_ptr(b), w, h, _ptr(depth), _ptr(conf))
    return depth, conf
