"""sgm_compute_device_noalloc through the C ABI: the entry point without a context or any device
allocation gives, output for output, the bytes sgm_compute_device gives on a context, on calls in a row
that share its static scratch (the second and third smaller than the first, so stale scratch from an
earlier call would show), and refuses pairs whose scratch exceeds the static array."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import speckle_pair
from test_gpu_parity import FIELDS, VOLUMES, assert_same

pytestmark = pytest.mark.gpu

BASE = dict(census_radius_x=4, census_radius_y=3, p1=10, p2=120)


def _lib():
    L = sgm.lib()
    L.sgm_compute_device_noalloc.argtypes = [C.POINTER(sgm.CSgmParams), C.c_void_p, C.c_void_p, C.c_ssize_t, C.c_int,
                                             C.c_int, C.POINTER(sgm.CSgmOutputs), C.c_void_p]
    L.sgm_compute_device_noalloc.restype = C.c_int
    L.sgm_noalloc_scratch_bytes.restype = C.c_size_t
    return L


def _outputs(w, h, nd, fields):
    bufs = {}
    for f in fields:
        shape = (h, w, nd) if f in sgm.VOLUME_FIELDS else (h, w)
        b = sgm.DeviceBuffer(int(np.prod(shape)) * np.dtype(sgm.OUTPUT_DTYPES[f]).itemsize)
        b.memset(0xEE)
        bufs[f] = (b, shape)
    return bufs


def _download(bufs):
    return {f: b.download(shape, sgm.OUTPUT_DTYPES[f]) for f, (b, shape) in bufs.items()}


def _free(bufs):
    for b, _ in bufs.values():
        b.free()


CASES = [  # (w, h, nd, eight_paths, seed, fields): each later call smaller than the first
    (640, 512, 64, False, 31, FIELDS + ["cost_volume"]),
    (320, 256, 32, True, 32, FIELDS + ["sum_volume"]),
    (250, 130, 48, True, 33, [f for f in FIELDS if f != "disparity"]),  # disparity only in the scratch
]


def test_noalloc_equals_the_context_entry_point():
    L = _lib()
    assert L.sgm_noalloc_scratch_bytes() == 128 << 20
    st = None
    for w, h, nd, eight, seed, fields in CASES:
        pair = speckle_pair(w, h, nd, seed)
        kw = dict(BASE, min_disparity=pair.min_disparity, num_disparities=nd, eight_paths=eight)
        print(f"noalloc {w}x{h}x{nd} eight_paths={eight} fields={fields}", flush=True)
        img = sgm.DeviceBuffer(2 * w * h)
        img.upload(pair.left, 0)
        img.upload(pair.right, w * h)
        mine, theirs = _outputs(w, h, nd, fields), _outputs(w, h, nd, FIELDS + [f for f in fields if f in VOLUMES])
        st = sgm.HipStereo.create(sgm.SgmParams(**kw), w, h)
        try:
            st.compute_device(img.ptr, img.ptr + w * h, w, w, h, {f: b.ptr for f, (b, _) in theirs.items()})
            co = sgm.CSgmOutputs()
            for f, (b, _) in mine.items():
                setattr(co, f, b.ptr)
            cp = sgm.SgmParams(**kw).to_c()
            rc = L.sgm_compute_device_noalloc(C.byref(cp), img.ptr, img.ptr + w * h, w, w, h, C.byref(co), st.stream())
            assert rc == 0, (L.sgm_last_error(None) or b"").decode()
            st.synchronize()
            got, ref = _download(mine), _download(theirs)
            assert_same(ref, got, list(mine), tag=f"noalloc vs sgm_compute_device {w}x{h}x{nd}")
            oref = O.sgm(pair.left, pair.right, O.SgmParams(**kw), volumes=True)
            assert_same(oref, got, list(mine), tag=f"noalloc vs oracle {w}x{h}x{nd}")
        finally:
            st.close()
            _free(mine)
            _free(theirs)
            img.free()


def test_noalloc_refuses_pairs_beyond_its_scratch():
    L = _lib()
    kw = dict(BASE, min_disparity=0, num_disparities=128, eight_paths=True)
    cp = sgm.SgmParams(**kw).to_c()
    co = sgm.CSgmOutputs()
    img = sgm.DeviceBuffer(2 * 1280 * 1024)  # real images: the refusal comes before any device work
    try:
        rc = L.sgm_compute_device_noalloc(C.byref(cp), img.ptr, img.ptr + 1280 * 1024, 1280, 1280, 1024, C.byref(co),
                                          None)
    finally:
        img.free()
    assert sgm.ERRC[rc] == "invalid_argument"
    assert b"scratch" in (L.sgm_last_error(None) or b"")
