"""The Python wrapper refuses pairs whose two images differ in size before any native call: the C
entry points take one width, height and stride for both images, so a smaller second image would be
read past its end on the host."""
import numpy as np
import pytest

from open_source_einstar_amd import sgm


def _no_gpu_stereo():
    st = sgm.HipStereo.__new__(sgm.HipStereo)
    st._h = None  # never reached: the checks raise first
    st.rect_size = (64, 48)
    return st


@pytest.mark.parametrize("call", ["compute", "compute_disparity", "compute_disparity_raw", "compute_frame_raw",
                                  "compute_raw"])
def test_mismatched_pair_is_refused(call):
    st = _no_gpu_stereo()
    a, b = np.zeros((48, 64), np.uint8), np.zeros((40, 64), np.uint8)
    with pytest.raises(sgm.SgmError) as e:
        getattr(st, call)(a, b)
    assert e.value.errc == "invalid_argument"


def test_mismatched_full_resolution_pair_is_refused():
    st = _no_gpu_stereo()
    a = np.zeros((48, 64), np.uint8)
    with pytest.raises(sgm.SgmError):
        st.compute_disparity(a, a, np.zeros((96, 128), np.uint8), np.zeros((96, 120), np.uint8))
    with pytest.raises(sgm.SgmError):
        st.compute_disparity(a, a, np.zeros((96, 128), np.uint8), None)
