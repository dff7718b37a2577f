"""MI355X-native (gfx950) SGM stereo matcher: census -> Hamming cost -> 4/8-path semi-global
aggregation -> WTA + uniqueness + subpixel -> left/right check -> 3x3 median, behind a C ABI
(include/sgm_hip.h) that replaces the reference's Metal stereo backend."""
from .sgm import (Batch, DeviceBuffer, HipStereo, INVALID_DISPARITY, NO_MATCH, PairSpec, SgmError, SgmParams,
                  STATUS_LR, STATUS_UNIQUE, STATUS_VALID, STATUS_WINDOW, mem_info)

__all__ = ["Batch", "DeviceBuffer", "HipStereo", "INVALID_DISPARITY", "NO_MATCH", "PairSpec", "SgmError",
           "SgmParams", "STATUS_LR", "STATUS_UNIQUE", "STATUS_VALID", "STATUS_WINDOW", "mem_info"]
