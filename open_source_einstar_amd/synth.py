"""Synthetic rectified IR speckle pairs (ctypes over libeinstar_synth.so; see csrc/speckle_synth.c)."""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libeinstar_synth.so")


class CSynthParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("num_disparities", C.c_int32),
        ("min_cells", C.c_int32), ("max_cells", C.c_int32), ("seed", C.c_uint64),
        ("z_min", C.c_double), ("z_max", C.c_double), ("max_gradient", C.c_double),
        ("dot_density", C.c_double), ("blur_sigma", C.c_double), ("noise_sigma", C.c_double),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m open_source_einstar_amd.build synth`")
        L = C.CDLL(LIB_PATH)
        L.synth_speckle_pair.argtypes = [C.POINTER(CSynthParams), C.c_void_p, C.c_void_p, C.c_void_p]
        L.synth_speckle_pair.restype = C.c_int32
        _lib = L
    return _lib


@dataclass
class SpecklePair:
    left: np.ndarray
    right: np.ndarray
    gt_disparity: np.ndarray
    min_disparity: int
    num_disparities: int


def speckle_pair(width: int, height: int, num_disparities: int, seed: int, *, min_cells: int = 12,
                 max_cells: int = 24, z_min: float = 200.0, z_max: float = 600.0, max_gradient: float = 0.15,
                 dot_density: float = 0.1, blur_sigma: float = 1.0, noise_sigma: float = 2.0) -> SpecklePair:
    """Piecewise-planar speckle scene at z in [z_min, z_max] mm whose disparities fall inside
    [min_d + 2, min_d + num_disparities - 2); returns the pair, ground truth and min_d."""
    p = CSynthParams(width, height, num_disparities, min_cells, max_cells, seed, z_min, z_max, max_gradient,
                     dot_density, blur_sigma, noise_sigma)
    left = np.empty((height, width), np.uint8)
    right = np.empty((height, width), np.uint8)
    gt = np.empty((height, width), np.float32)
    md = _load().synth_speckle_pair(C.byref(p), left.ctypes.data, right.ctypes.data, gt.ctypes.data)
    if md == -(2**31):
        raise ValueError("synth_speckle_pair: invalid parameters")
    return SpecklePair(left, right, gt, int(md), num_disparities)


@dataclass
class RemapPair:
    """Left and right remap tables (calib::RemapTable, libs/calib/include/einstar/calib/rectify.hpp:40-43):
    for each rectified pixel, the float32 source coordinates in the raw (distorted) frame."""

    left_x: np.ndarray
    left_y: np.ndarray
    right_x: np.ndarray
    right_y: np.ndarray


def _rotation(rng, max_angle_deg: float) -> np.ndarray:
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    a = np.deg2rad(rng.uniform(-max_angle_deg, max_angle_deg))
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(a) * k + (1 - np.cos(a)) * (k @ k)


def rectification_maps(map_w: int, map_h: int, raw_w: int, raw_h: int, seed: int, *, max_angle_deg: float = 0.6,
                       k1: float = -0.08, k2: float = 0.02, tangential: float = 5e-4) -> RemapPair:
    """Rectification tables of a synthetic rig, built like calib::build_remap (rectify.cpp:117-131):
    a rectified pinhole pixel -> ray, rotated back into the raw camera by a small random rotation,
    projected through a Brown radial/tangential model with a jittered principal point, and stored as
    float32. Pixels near the border may map outside the raw frame (rectification then gives 0)."""
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:map_h, 0:map_w].astype(np.float64)
    out = []
    for _ in range(2):
        f = 0.9 * raw_w * rng.uniform(0.98, 1.02)
        cx = (raw_w - 1) / 2 + rng.uniform(-0.01, 0.01) * raw_w
        cy = (raw_h - 1) / 2 + rng.uniform(-0.01, 0.01) * raw_h
        fr = 0.9 * raw_w * map_w / raw_w
        rays = np.stack([(xs - (map_w - 1) / 2) / fr, (ys - (map_h - 1) / 2) / fr, np.ones_like(xs)], axis=-1)
        r = rays @ _rotation(rng, max_angle_deg)  # row vectors: (R^T ray)^T = ray^T R
        nx, ny = r[..., 0] / r[..., 2], r[..., 1] / r[..., 2]
        r2 = nx * nx + ny * ny
        radial = 1 + k1 * r2 + k2 * r2 * r2
        t1, t2 = tangential * rng.uniform(-1, 1), tangential * rng.uniform(-1, 1)
        dx = nx * radial + 2 * t1 * nx * ny + t2 * (r2 + 2 * nx * nx)
        dy = ny * radial + t1 * (r2 + 2 * ny * ny) + 2 * t2 * nx * ny
        out += [(f * dx + cx).astype(np.float32), (f * dy + cy).astype(np.float32)]
    return RemapPair(*out)
