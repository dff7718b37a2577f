"""Pins of the oracle's census transform and cost fill (stereo.cpp:328-346, 366-380).

Known answers fix the bit order (rows outer, columns inner, first compare in the top bit), the
border rule and the image-bound cases of the cost fill; they do not restate the oracle's loops.
"""
import numpy as np
import pytest

import oracle as O


def window_image(rx, ry, dark):
    """(2rx+1+2) x (2ry+1+2) image, centre 100, window 200, the offsets in `dark` set to 50."""
    w, h = 2 * rx + 3, 2 * ry + 3
    img = np.full((h, w), 200, np.uint8)
    cx, cy = rx + 1, ry + 1
    img[cy, cx] = 100
    for (u, v) in dark:
        img[cy + v, cx + u] = 50
    return img, cx, cy


@pytest.mark.parametrize("rx,ry", [(4, 3), (2, 2), (1, 1), (3, 1)])
def test_census_bit_order_known_answers(rx, ry):
    n = (2 * rx + 1) * (2 * ry + 1) - 1
    # First compare (top-left) lands in bit n-1, the last (bottom-right) in bit 0.
    img, cx, cy = window_image(rx, ry, [(-rx, -ry)])
    assert O.census(img, rx, ry)[cy, cx] == np.uint64(1) << np.uint64(n - 1)
    img, cx, cy = window_image(rx, ry, [(rx, ry)])
    assert O.census(img, rx, ry)[cy, cx] == 1
    # Second row's first pixel is compare number 2rx+1 (row-major scan): a column-major scan
    # would put it at compare 1.
    img, cx, cy = window_image(rx, ry, [(-rx, -ry + 1)])
    assert O.census(img, rx, ry)[cy, cx] == np.uint64(1) << np.uint64(n - 1 - (2 * rx + 1))
    # The pixel right after the centre is compare n/2 (the centre itself is skipped).
    img, cx, cy = window_image(rx, ry, [(1, 0)])
    assert O.census(img, rx, ry)[cy, cx] == np.uint64(1) << np.uint64(n - 1 - n // 2)


def test_census_9x7_full_word():
    img, cx, cy = window_image(4, 3, [(u, v) for v in range(-3, 4) for u in range(-4, 5) if (u, v) != (0, 0)])
    assert O.census(img, 4, 3)[cy, cx] == (1 << 62) - 1


def test_census_equal_is_not_darker_and_border_is_zero():
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (20, 30), dtype=np.uint8)
    c = O.census(img, 4, 3)
    assert not c[:3].any() and not c[-3:].any() and not c[:, :4].any() and not c[:, -4:].any()
    assert not O.census(np.full((12, 12), 77, np.uint8), 2, 2).any()
    # Popcount = number of strictly darker neighbours (order-free check of the content).
    for y in range(3, 17):
        for x in range(4, 26):
            win = img[y - 3:y + 4, x - 4:x + 5].astype(int)
            assert bin(int(c[y, x])).count("1") == int((win < int(img[y, x])).sum())


def test_census_tiny_images():
    for (h, w) in [(1, 1), (1, 9), (7, 1), (6, 8)]:
        c = O.census(np.arange(w * h, dtype=np.uint8).reshape(h, w), 4, 3)
        assert c.shape == (h, w) and not c.any()


def test_cost_fill_bounds_and_popcount():
    h, w, nd = 2, 8, 6
    cl = np.zeros((h, w), np.uint64)
    cr = np.zeros((h, w), np.uint64)
    cl[0, 5] = np.uint64((1 << 62) - 1)
    cr[0, :] = np.uint64(0b1011)  # popcount 3 in every right word
    invalid = 63
    for min_d in (0, 2, -3):
        c = O.cost_volume(cl, cr, min_d, nd, invalid)
        for x in range(w):
            for k in range(nd):
                xr = x - (min_d + k)
                if xr < 0 or xr >= w:
                    assert c[0, x, k] == invalid, (min_d, x, k)
                elif x == 5:
                    assert c[0, x, k] == 62 - 3
                else:
                    assert c[0, x, k] == 3
    # x = 0, min_d = 0: only k = 0 has a right pixel; the rest stay at the invalid fill.
    c = O.cost_volume(cl, cr, 0, nd, invalid)
    assert c[1, 0, 0] == 0 and (c[1, 0, 1:] == invalid).all()
