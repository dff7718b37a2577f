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
