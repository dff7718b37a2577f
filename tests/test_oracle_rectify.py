"""Pins of the oracle's remap (libs/calib/src/rectify.cpp:133-148) and downsample2
(libs/depth/src/stereo.cpp:348-359) by known answers and exactness of bilinear sampling on linear
images."""
import numpy as np

import oracle as O


def test_identity_map_copies_except_the_last_row_and_column():
    rng = np.random.default_rng(0)
    src = rng.integers(0, 256, (9, 12), dtype=np.uint8)
    ys, xs = np.mgrid[0:9, 0:12].astype(np.float32)
    out = O.remap(src, xs, ys)
    # x0 + 1 >= width or y0 + 1 >= height counts as outside (the 2x2 neighbourhood must fit).
    np.testing.assert_array_equal(out[:-1, :-1], src[:-1, :-1])
    assert not out[-1].any() and not out[:, -1].any()


def test_integer_and_half_pixel_shifts():
    src = np.zeros((6, 8), np.uint8)
    src[2, 3], src[2, 4] = 10, 13
    ys, xs = np.mgrid[0:6, 0:8].astype(np.float32)
    out = O.remap(src, xs + 1, ys)
    assert out[2, 2] == 10 and out[2, 3] == 13
    out = O.remap(src, xs + 0.5, ys)
    assert out[2, 3] == 12  # (10 + 13) / 2 = 11.5, + 0.5 -> 12.0, truncated
    assert out[2, 2] == 5   # (0 + 10) / 2 = 5.0, + 0.5 -> 5.5, truncated to 5


def test_out_of_bounds_and_negative_coordinates_are_zero():
    src = np.full((5, 5), 200, np.uint8)
    xs = np.array([[-0.5, -0.001, 0.0, 3.0, 3.5, 4.0]], np.float32)
    ys = np.zeros_like(xs) + 1.0
    out = O.remap(src, xs, ys)
    np.testing.assert_array_equal(out, [[0, 0, 200, 200, 200, 0]])


def test_bilinear_is_exact_on_linear_images():
    # src(x, y) = 2x + 3y: bilinear sampling reproduces 2 sx + 3 sy up to float rounding, then
    # +0.5 and truncation round it to the nearest integer (ties up).
    ys, xs = np.mgrid[0:40, 0:50]
    src = (2 * xs + 3 * ys).astype(np.uint8)
    rng = np.random.default_rng(1)
    mx = rng.uniform(0, 48, (30, 30)).astype(np.float32)
    my = rng.uniform(0, 38, (30, 30)).astype(np.float32)
    out = O.remap(src, mx, my).astype(np.int64)
    exact = 2 * mx.astype(np.float64) + 3 * my.astype(np.float64)
    assert np.abs(out - np.floor(exact + 0.5)).max() <= 1
    assert (out == np.floor(exact + 0.5)).mean() > 0.99


def test_downsample_rounding_known_answers():
    blocks = {(0, 0, 0, 1): 0, (1, 1, 1, 1): 1, (0, 1, 1, 1): 1, (0, 0, 1, 1): 1, (254, 255, 255, 255): 255,
              (0, 0, 0, 2): 1, (3, 3, 3, 4): 3}
    for (a, b, c, d), want in blocks.items():
        img = np.array([[a, b], [c, d]], np.uint8)
        assert O.downsample2(img)[0, 0] == want, (a, b, c, d)
    odd = np.arange(35, dtype=np.uint8).reshape(5, 7)
    assert O.downsample2(odd).shape == (2, 3)


def test_rectify_half_is_remap_then_downsample():
    rng = np.random.default_rng(2)
    raw = rng.integers(0, 256, (64, 80), dtype=np.uint8)
    ys, xs = np.mgrid[0:48, 0:60].astype(np.float32)
    mx, my = xs * 1.1 + 3.3, ys * 0.97 + 2.2
    np.testing.assert_array_equal(O.rectify_half(raw, mx, my), O.downsample2(O.remap(raw, mx, my)))
