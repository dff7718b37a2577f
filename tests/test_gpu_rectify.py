"""Fused rectification (remap at the maps' resolution + downsample2 in one kernel, maps resident on
the device) against the oracle's calib::remap + downsample2 (libs/calib/src/rectify.cpp:133-148,
libs/depth/src/stereo.cpp:348-359): the rectified pair is bit-exact, as in the reference's own
test (tests/depth/test_metal_stereo.cpp:96-103), and so is everything SGM computes from it."""
import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import rectification_maps, speckle_pair
from test_gpu_parity import FIELDS, assert_same

pytestmark = pytest.mark.gpu


def raw_pair(raw_w, raw_h, seed):
    p = speckle_pair(raw_w, raw_h, 16, seed)
    return p.left, p.right


def check_rectified(st, maps, raw_l, raw_r, params, tag):
    got = st.compute_raw(raw_l, raw_r, outputs=FIELDS, rectified=True)
    rl = O.rectify_half(raw_l, maps.left_x, maps.left_y)
    rr = O.rectify_half(raw_r, maps.right_x, maps.right_y)
    np.testing.assert_array_equal(got["rect_left"], rl, err_msg=f"{tag} left")
    np.testing.assert_array_equal(got["rect_right"], rr, err_msg=f"{tag} right")
    ref = O.sgm(rl, rr, O.SgmParams(**params))
    assert_same(ref, got, FIELDS, tag=tag)
    return got


CASES = [
    # (map_w, map_h, raw_w, raw_h, nd, seed)
    (128, 96, 128, 96, 16, 1),
    (130, 98, 120, 90, 24, 2),     # raw smaller than the maps: more samples fall outside
    (257, 131, 300, 140, 32, 3),   # odd map sizes: downsample2 drops the last column and row
    (640, 512, 700, 560, 64, 4),
]


@pytest.mark.parametrize("mw,mh,rw,rh,nd,seed", CASES, ids=[f"{c[0]}x{c[1]}-raw{c[2]}x{c[3]}" for c in CASES])
def test_rectified_pair_and_disparity_bit_exact(mw, mh, rw, rh, nd, seed):
    print(f"rectify map {mw}x{mh} raw {rw}x{rh} D={nd} seed={seed}", flush=True)
    maps = rectification_maps(mw, mh, rw, rh, seed)
    raw_l, raw_r = raw_pair(rw, rh, 100 + seed)
    params = dict(min_disparity=-4, num_disparities=nd, eight_paths=True, census_radius_x=4, census_radius_y=3,
                  p1=10, p2=120)
    with sgm.HipStereo.create(sgm.SgmParams(**params), mw // 2, mh // 2) as st:
        st.set_rectification(maps.left_x, maps.left_y, maps.right_x, maps.right_y, rw, rh)
        check_rectified(st, maps, raw_l, raw_r, params, f"map{mw}x{mh}")
        t = st.last_timings()
        assert t["rectify_ms"] > 0 and t["gpu_ms"] >= t["rectify_ms"]


def test_boundary_coordinates():
    """Exact borders of the 2x2 neighbourhood, -0.0, values just inside, far outside and large
    finite coordinates, on both axes."""
    rw, rh = 40, 30
    edge = np.array([0.0, -0.0, -1e-6, 1e-6, rw - 1.0, rw - 1.0 - 1e-4, rw - 2.0, rw - 2.5, 1e6, -1e6, 17.5, 3.25,
                     np.nextafter(np.float32(rw - 1), np.float32(0)), rw + 3.0, 0.999999, 38.99], np.float32)
    mw, mh = 16, 16
    mx = np.tile(edge, (mh, 1))
    my = np.tile(np.linspace(-1, rh + 1, mh, dtype=np.float32)[:, None], (1, mw))
    my[::3] = np.float32(rh - 1)
    my[1::5] = np.float32(rh - 2) + np.float32(0.75)
    rng = np.random.default_rng(7)
    raw_l = rng.integers(0, 256, (rh, rw), dtype=np.uint8)
    raw_r = rng.integers(0, 256, (rh, rw), dtype=np.uint8)
    raw_l[:, -2:] = 255  # bright border: rounding at 255 and the clamp
    params = dict(num_disparities=8)
    with sgm.HipStereo.create(sgm.SgmParams(**params), mw // 2, mh // 2) as st:
        st.set_rectification(mx, my, my.T.copy() * (rw / rh), mx.T.copy() * (rh / rw), rw, rh)
        got = st.compute_raw(raw_l, raw_r, rectified=True)
    np.testing.assert_array_equal(got["rect_left"], O.rectify_half(raw_l, mx, my))
    np.testing.assert_array_equal(got["rect_right"], O.rectify_half(raw_r, my.T.copy() * (rw / rh), mx.T.copy() * (rh / rw)))


@pytest.mark.parametrize("rw,rh", [(1, 30), (40, 1), (1, 1), (2, 2)])
def test_one_pixel_raw_frames(rw, rh):
    """Raw frames one pixel wide or tall have no 2x2 neighbourhood: every sample is outside and the
    rectified images are 0 (a 2x2 frame has one); the kernel reads nothing outside the frame."""
    mw, mh = 16, 12
    rng = np.random.default_rng(rw * 100 + rh)
    mx = rng.uniform(-1, 3, (mh, mw)).astype(np.float32)
    my = rng.uniform(-1, 3, (mh, mw)).astype(np.float32)
    raw_l = rng.integers(0, 256, (rh, rw), dtype=np.uint8)
    raw_r = rng.integers(0, 256, (rh, rw), dtype=np.uint8)
    with sgm.HipStereo.create(sgm.SgmParams(num_disparities=8), mw // 2, mh // 2) as st:
        st.set_rectification(mx, my, my, mx, rw, rh)
        got = st.compute_raw(raw_l, raw_r, rectified=True)
    np.testing.assert_array_equal(got["rect_left"], O.rectify_half(raw_l, mx, my))
    np.testing.assert_array_equal(got["rect_right"], O.rectify_half(raw_r, my, mx))
    if rw < 2 or rh < 2:
        assert not got["rect_left"].any() and not got["rect_right"].any()


def test_full_size_1280x1024x128_from_2560x2048_raw():
    mw, mh, nd = 2560, 2048, 128
    maps = rectification_maps(mw, mh, mw, mh, 11)
    raw_l, raw_r = raw_pair(mw, mh, 12)
    params = dict(min_disparity=-8, num_disparities=nd, eight_paths=True, census_radius_x=4, census_radius_y=3,
                  p1=10, p2=120)
    print("rectify 2560x2048 -> 1280x1024x128", flush=True)
    with sgm.HipStereo.create(sgm.SgmParams(**params), mw // 2, mh // 2) as st:
        st.set_rectification(maps.left_x, maps.left_y, maps.right_x, maps.right_y, mw, mh)
        a = check_rectified(st, maps, raw_l, raw_r, params, "full")
        # A second frame through the same resident maps gives the same bytes.
        b = st.compute_raw(raw_l, raw_r, outputs=FIELDS, rectified=True)
        assert_same(a, b, FIELDS + ["rect_left", "rect_right"], tag="repeat")


def test_device_entry_point_and_strided_raw_frames():
    mw, mh, rw, rh, nd = 200, 160, 210, 170, 32
    maps = rectification_maps(mw, mh, rw, rh, 21)
    raw_l, raw_r = raw_pair(rw, rh, 22)
    pitch = rw + 37
    params = sgm.SgmParams(min_disparity=-2, num_disparities=nd, eight_paths=True)
    w, h = mw // 2, mh // 2
    with sgm.HipStereo.create(params, w, h) as st:
        st.set_rectification(maps.left_x, maps.left_y, maps.right_x, maps.right_y, rw, rh)
        host = st.compute_raw(raw_l, raw_r, outputs=FIELDS, rectified=True)
        padded = np.zeros((2 * rh, pitch), np.uint8)
        padded[:rh, :rw], padded[rh:, :rw] = raw_l, raw_r
        raw = sgm.DeviceBuffer(padded.size)
        raw.upload(padded, 0)
        rect = sgm.DeviceBuffer(2 * w * h)
        outs = {f: sgm.DeviceBuffer(w * h * np.dtype(sgm.OUTPUT_DTYPES[f]).itemsize) for f in FIELDS}
        st.compute_raw_device(raw.ptr, raw.ptr + rh * pitch, pitch, rw, rh, {f: b.ptr for f, b in outs.items()},
                              rect.ptr, rect.ptr + w * h)
        st.synchronize()
        dev = {f: b.download((h, w), sgm.OUTPUT_DTYPES[f]) for f, b in outs.items()}
        both = rect.download((2 * h, w), np.uint8)
        dev["rect_left"], dev["rect_right"] = both[:h], both[h:]
    assert_same(host, dev, FIELDS + ["rect_left", "rect_right"], tag="device")


def test_errors_follow_the_reference():
    with sgm.HipStereo.create(sgm.SgmParams(num_disparities=16), 64, 48) as st:
        raw = np.zeros((100, 130), np.uint8)
        with pytest.raises(sgm.SgmError) as e:
            st.compute_raw(raw, raw)
        assert e.value.errc == "invalid_argument" and "set_rectification() not called" in str(e.value)
        big = np.zeros((100, 130), np.float32)  # stereo size 65 x 50 exceeds the context's 64 x 48
        with pytest.raises(sgm.SgmError) as e:
            st.set_rectification(big, big, big, big, 130, 100)
        assert "twice the stereo resolution" in str(e.value)
        ok = np.zeros((96, 128), np.float32)
        st.set_rectification(ok, ok, ok, ok, 130, 100)
        with pytest.raises(sgm.SgmError) as e:
            st.compute_raw(np.zeros((101, 130), np.uint8), np.zeros((101, 130), np.uint8))
        assert "raw size mismatch" in str(e.value)
        got = st.compute_raw(raw, raw, rectified=True)  # every sample at (0, 0) is inside: copies raw[0, 0]
        assert got["rect_left"].shape == (48, 64) and not got["rect_left"].any()


def remap_sweep(n, seed):
    """Seeded random maps: smooth rectification-like tables mixed with uniform noise over and past
    the raw frame, exact integer and half-pixel coordinates, -0.0 and values just inside the last
    row/column, at random map and raw sizes."""
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        mw, mh = 2 * int(rng.integers(4, 161)), 2 * int(rng.integers(4, 121))
        rw, rh = int(rng.integers(2, 341)), int(rng.integers(2, 261))
        cases.append((mw, mh, rw, rh, int(rng.integers(0, 2**31))))
    return cases


def random_map(mw, mh, rw, rh, rng):
    ys, xs = np.mgrid[0:mh, 0:mw].astype(np.float32)
    mx = xs * np.float32(rw / mw) + rng.normal(0, 2, (mh, mw)).astype(np.float32)
    my = ys * np.float32(rh / mh) + rng.normal(0, 2, (mh, mw)).astype(np.float32)
    noise = rng.random((mh, mw)) < 0.2
    mx[noise] = rng.uniform(-3, rw + 3, int(noise.sum())).astype(np.float32)
    my[noise] = rng.uniform(-3, rh + 3, int(noise.sum())).astype(np.float32)
    special = np.array([0.0, -0.0, 0.5, 1.0, 1.5, rw - 2.0, rw - 1.5, rw - 1.0,
                        np.nextafter(np.float32(rw - 1), np.float32(0))], np.float32)
    pick = rng.random((mh, mw)) < 0.1
    mx[pick] = rng.choice(special, int(pick.sum()))
    special_y = np.array([0.0, -0.0, 0.5, rh - 2.0, rh - 1.5, rh - 1.0,
                          np.nextafter(np.float32(rh - 1), np.float32(0))], np.float32)
    pick = rng.random((mh, mw)) < 0.1
    my[pick] = rng.choice(special_y, int(pick.sum()))
    return mx, my


REMAP_SWEEP = remap_sweep(16, 404)


@pytest.mark.parametrize("mw,mh,rw,rh,seed", REMAP_SWEEP, ids=[f"remap{i}-{c[0]}x{c[1]}-raw{c[2]}x{c[3]}"
                                                               for i, c in enumerate(REMAP_SWEEP)])
def test_random_maps_bit_exact(mw, mh, rw, rh, seed):
    rng = np.random.default_rng(seed)
    lx, ly = random_map(mw, mh, rw, rh, rng)
    rx, ry = random_map(mw, mh, rw, rh, rng)
    raw_l = rng.integers(0, 256, (rh, rw), dtype=np.uint8)
    raw_r = rng.integers(0, 256, (rh, rw), dtype=np.uint8)
    print(f"random maps {mw}x{mh} raw {rw}x{rh} seed={seed}", flush=True)
    with sgm.HipStereo.create(sgm.SgmParams(num_disparities=16), mw // 2, mh // 2) as st:
        st.set_rectification(lx, ly, rx, ry, rw, rh)
        got = st.compute_raw(raw_l, raw_r, rectified=True)
    np.testing.assert_array_equal(got["rect_left"], O.rectify_half(raw_l, lx, ly))
    np.testing.assert_array_equal(got["rect_right"], O.rectify_half(raw_r, rx, ry))
