"""Pins of the oracle's WTA / uniqueness / subpixel (stereo.cpp:92-96, 411-448), left/right check
(:450-461) and median3 (:129-147) by crafted inputs, textbook closed forms and brute force."""
import numpy as np
import pytest

import oracle as O

INV = np.float32(-1e9)


def one_pixel_S(vals, w=64, x=63, min_d=0):
    """A 1-row sum volume whose pixel x holds `vals`, every other pixel a flat 1000."""
    nd = len(vals)
    S = np.full((1, w, nd), 1000, np.uint16)
    S[0, x] = vals
    return S


def test_first_strict_argmin_and_second_excludes_neighbours():
    p = O.SgmParams(num_disparities=8, uniqueness=0.9, subpixel=False)
    o = O.wta(one_pixel_S([50, 40, 40, 45, 60, 39, 39, 70]), p)
    # Ties go to the lowest k: best = 5 (39), second = min over |k-5| > 1 = 40.
    assert o["disparity_int"][0, 63] == 5
    assert o["best_cost"][0, 63] == 39 and o["second_cost"][0, 63] == 40
    assert o["disparity"][0, 63] == INV  # 39 >= 0.9 * 40 = 36: not unique
    assert o["status"][0, 63] == O.ST_WINDOW
    o = O.wta(one_pixel_S([90, 30, 100, 100, 100, 100, 100, 100]), p)
    assert o["disparity"][0, 63] == 1.0 and o["status"][0, 63] == O.ST_WINDOW | O.ST_UNIQUE


def test_uniqueness_boundary_is_inclusive():
    p = O.SgmParams(num_disparities=6, uniqueness=0.9, subpixel=False)
    # best 90, second 100: 90 >= 0.9f * 100 -> rejected; 89 accepted.
    o = O.wta(one_pixel_S([90, 200, 200, 100, 200, 200]), p)
    assert o["disparity"][0, 63] == INV
    o = O.wta(one_pixel_S([89, 200, 200, 100, 200, 200]), p)
    assert o["disparity"][0, 63] == 0.0
    # Only adjacent candidates: no second, no uniqueness test.
    o = O.wta(one_pixel_S([100, 99]), O.SgmParams(num_disparities=2, uniqueness=0.9, subpixel=False))
    assert o["second_cost"][0, 63] == 0xFFFF and o["disparity"][0, 63] == 1.0


@pytest.mark.parametrize("vertex", [2.25, 1.6, 2.5, 2.0])
def test_subpixel_is_the_parabola_vertex(vertex):
    # Integer samples of 16 (k - vertex)^2 + 4 around k = 2: the offset is the vertex of the
    # interpolating parabola (numpy.polyfit), clamped to +-0.5.
    ks = np.array([1, 2, 3])
    vals = np.rint(16 * (ks - vertex) ** 2 + 4).astype(int)
    S = [500, *vals, 500, 500]
    p = O.SgmParams(num_disparities=len(S), uniqueness=1.0, subpixel=True)
    o = O.wta(one_pixel_S(S), p)
    a, b, _ = np.polyfit(ks.astype(float), vals.astype(float), 2)
    expect = float(np.clip(-b / (2 * a) - 2.0, -0.5, 0.5))
    assert o["disparity_int"][0, 63] == 2
    assert abs(float(o["disparity"][0, 63]) - (2.0 + expect)) < 1e-6
    assert O.subpixel_offset(10.0, 10.0, 10.0) == 0.0           # flat: denominator 0
    assert O.subpixel_offset(0.0, 10.0, 0.0) == 0.0             # maximum: denominator < 0
    assert O.subpixel_offset(0.0, 10.0, 100.0) == -0.5          # -0.625 clamped


def test_no_subpixel_at_window_edges():
    p = O.SgmParams(num_disparities=5, uniqueness=1.0, subpixel=True)
    o = O.wta(one_pixel_S([10, 50, 60, 70, 80]), p)  # best at k = 0 = min_k
    assert o["disparity"][0, 63] == 0.0
    o = O.wta(one_pixel_S([80, 70, 60, 50, 10]), p)  # best at max_k - 1
    assert o["disparity"][0, 63] == 4.0


def test_candidate_window_at_image_borders():
    # x - (min_d + k) must lie in [0, w): at x = 1 with min_d = 0 only k in {0, 1}.
    nd, w = 6, 10
    S = np.full((1, w, nd), 100, np.uint16)
    S[0, 1] = [50, 40, 1, 1, 1, 1]
    p = O.SgmParams(num_disparities=nd, uniqueness=1.0, subpixel=False)
    o = O.wta(S, p)
    assert o["disparity_int"][0, 1] == 1 and o["best_cost"][0, 1] == 40
    # Negative min_d: at x = w-1, candidates need x - min_d - k <= w-1 -> k >= -min_d.
    p = O.SgmParams(min_disparity=-3, num_disparities=nd, uniqueness=1.0, subpixel=False)
    S2 = np.full((1, w, nd), 100, np.uint16)
    S2[0, w - 1] = [1, 1, 1, 30, 20, 50]
    o = O.wta(S2, p)
    assert o["disparity_int"][0, w - 1] == -3 + 4
    # Empty window: x - min_d + 1 <= 0.
    p = O.SgmParams(min_disparity=5, num_disparities=nd)
    o = O.wta(np.full((1, w, nd), 7, np.uint16), p)
    assert (o["disparity_int"][0, :5] == O.NO_MATCH).all() and (o["status"][0, :5] == 0).all()


def test_right_wta_first_argmin_over_the_diagonal():
    nd, w = 4, 6
    rng = np.random.default_rng(1)
    S = rng.integers(1, 50, (1, w, nd)).astype(np.uint16)
    S[0, 3, 1] = 0  # xr = 3 - 0 - 1 = 2
    S[0, 4, 2] = 0  # also xr = 2, tie: the lower k wins
    p = O.SgmParams(num_disparities=nd, uniqueness=1.0)
    o = O.wta(S, p)
    assert o["right_disparity"][0, 2] == 1
    for xr in range(w):  # brute force
        cands = [(int(S[0, xr + k, k]), k) for k in range(nd) if xr + k < w]
        assert o["right_disparity"][0, xr] == min(cands)[1]


def test_lr_check_rounds_halves_away_from_zero():
    w = 8
    rd = np.full((1, w), O.NO_MATCH, np.int32)
    rd[0, 3] = 2
    rd[0, 2] = 10
    # x = 5, d = 2.5: x - d = 2.5 -> lround 3 (passes: |2 - 2.5| <= 1); rint would pick 2 (fails).
    d = np.full((1, w), INV, np.float32)
    d[0, 5] = 2.5
    # x = 0, d = 0.5: x - d = -0.5 -> lround -1: outside the image (rint would give 0).
    rd[0, 0] = 0
    d[0, 0] = 0.5
    st = np.full((1, w), 3, np.uint8)
    out, st2 = O.lr_check(d, rd, 1, st)
    assert out[0, 5] == 2.5 and st2[0, 5] == 7
    assert out[0, 0] == INV and st2[0, 0] == 3
    # Tolerance is inclusive: |right - d| == lr_max_diff passes.
    d2 = np.full((1, w), INV, np.float32)
    d2[0, 6] = 3.0
    rd2 = np.full((1, w), 4, np.int32)
    assert O.lr_check(d2, rd2, 1)[0, 6] == 3.0
    rd2[:] = 5
    assert O.lr_check(d2, rd2, 1)[0, 6] == INV


def test_median3_brute_force():
    rng = np.random.default_rng(4)
    d = rng.integers(0, 40, (17, 23)).astype(np.float32) + rng.random((17, 23), dtype=np.float32)
    d[rng.random(d.shape) < 0.3] = INV
    m = O.median3(d)
    h, w = d.shape
    for y in range(h):
        for x in range(w):
            if d[y, x] < -1e8:
                assert m[y, x] == INV
                continue
            win = d[max(0, y - 1):y + 2, max(0, x - 1):x + 2].ravel()
            vals = np.sort(win[win > -1e8])
            assert m[y, x] == vals[len(vals) // 2]
