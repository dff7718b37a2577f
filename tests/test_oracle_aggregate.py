"""Pins of the oracle's path aggregation (stereo.cpp:44-89, 385-409) by exhaustive enumeration.

For each direction and pixel p the path is rebuilt by walking back from p until it leaves the
image (independent of the oracle's start-pixel list). Over every disparity sequence d_0..d_n on it,
    E_n(d) = min  sum_i C_i(d_i) + sum_i pen_i(d_{i-1}, d_i),
with pen = 0 (same), min(P1, p2_i) (|step| 1, or 2 with slant steps), p2_i (larger), p2_i adaptive
from the guide intensities. The recurrence of the reference subtracts the previous step's minimum,
so L_n(d) = E_n(d) - min E_{n-1} (L_0 = C_0); S = sum over directions.
"""
import itertools

import numpy as np
import pytest

import oracle as O

DIRS8 = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (1, -1), (-1, 1)]


def p2_step(p, i_now, i_prev):
    if p.adaptive_p2:
        return max(p.p1 + 1, p.p2 // (abs(int(i_now) - int(i_prev)) // 8 + 1))
    return p.p2


def brute_L(cost, guide, p, path):
    """L at the last pixel of `path` by enumerating every disparity sequence."""
    nd = cost.shape[2]
    n = len(path)
    seqs = np.array(list(itertools.product(range(nd), repeat=n)), dtype=np.int64)  # (nd^n, n)
    tot = np.zeros(len(seqs), np.int64)
    for i, (x, y) in enumerate(path):
        tot += cost[y, x, seqs[:, i]].astype(np.int64)
        if i:
            p2 = p2_step(p, guide[y, x], guide[path[i - 1][1], path[i - 1][0]])
            jump = np.abs(seqs[:, i] - seqs[:, i - 1])
            near = (jump == 1) | ((jump == 2) if p.slant_steps else False)
            pen = np.where(jump == 0, 0, np.where(near, min(p.p1, p2), p2))
            tot += pen

    def E(m):  # min over sequences of the first m pixels, as a function of d_{m-1}
        sub = np.zeros(nd ** m, np.int64)
        s = np.array(list(itertools.product(range(nd), repeat=m)), dtype=np.int64)
        for i in range(m):
            x, y = path[i]
            sub += cost[y, x, s[:, i]].astype(np.int64)
            if i:
                p2 = p2_step(p, guide[y, x], guide[path[i - 1][1], path[i - 1][0]])
                jump = np.abs(s[:, i] - s[:, i - 1])
                near = (jump == 1) | ((jump == 2) if p.slant_steps else False)
                sub += np.where(jump == 0, 0, np.where(near, min(p.p1, p2), p2))
        out = np.full(nd, np.iinfo(np.int64).max)
        np.minimum.at(out, s[:, m - 1], sub)
        return out

    En = np.full(nd, np.iinfo(np.int64).max)
    np.minimum.at(En, seqs[:, n - 1], tot)
    if n == 1:
        return En
    return En - E(n - 1).min()


def path_to(x, y, dx, dy, w, h):
    pts = []
    while 0 <= x < w and 0 <= y < h:
        pts.append((x, y))
        x -= dx
        y -= dy
    return pts[::-1]


def brute_S(cost, guide, p):
    h, w, nd = cost.shape
    dirs = DIRS8 if p.eight_paths else DIRS8[:4]
    S = np.zeros((h, w, nd), np.int64)
    for (dx, dy) in dirs:
        for y in range(h):
            for x in range(w):
                S[y, x] += brute_L(cost, guide, p, path_to(x, y, dx, dy, w, h))
    return S


CASES = [
    dict(p1=3, p2=11, adaptive_p2=True, slant_steps=True, eight_paths=True),
    dict(p1=3, p2=11, adaptive_p2=True, slant_steps=False, eight_paths=True),
    dict(p1=2, p2=9, adaptive_p2=False, slant_steps=True, eight_paths=False),
    dict(p1=7, p2=4, adaptive_p2=False, slant_steps=True, eight_paths=True),   # P2 < P1
    dict(p1=0, p2=40, adaptive_p2=True, slant_steps=True, eight_paths=True),
]


@pytest.mark.parametrize("case", CASES)
def test_aggregate_matches_exhaustive_enumeration(case):
    rng = np.random.default_rng(sum(case.values()) + 17)
    h, w, nd = 4, 5, 4
    cost = rng.integers(0, 12, (h, w, nd), dtype=np.uint16)
    guide = rng.integers(0, 256, (h, w), dtype=np.uint8)
    p = O.SgmParams(num_disparities=nd, **case)
    S = O.aggregate(cost, guide, p, threads=1)
    np.testing.assert_array_equal(S.astype(np.int64), brute_S(cost, guide, p))


def test_aggregate_single_direction_equals_enumeration_on_longer_line():
    rng = np.random.default_rng(3)
    h, w, nd = 1, 7, 3
    cost = rng.integers(0, 20, (h, w, nd), dtype=np.uint16)
    guide = rng.integers(0, 256, (h, w), dtype=np.uint8)
    p = O.SgmParams(num_disparities=nd, p1=4, p2=15)
    for dx in (1, -1):
        L = O.aggregate_direction(cost, guide, p, dx, 0)
        for x in range(w):
            np.testing.assert_array_equal(L[0, x].astype(np.int64), brute_L(cost, guide, p, path_to(x, 0, dx, 0, w, h)))


def test_zero_penalties_give_cost_times_paths():
    """P1 = P2 = 0 without adaptation: every step's best equals the previous minimum, so L = C."""
    rng = np.random.default_rng(5)
    cost = rng.integers(0, 63, (9, 11, 7), dtype=np.uint16)
    guide = rng.integers(0, 256, (9, 11), dtype=np.uint8)
    for eight in (False, True):
        p = O.SgmParams(num_disparities=7, p1=0, p2=0, adaptive_p2=False, eight_paths=eight)
        S = O.aggregate(cost, guide, p)
        np.testing.assert_array_equal(S, cost * (8 if eight else 4))


def test_every_pixel_covered_once_per_direction():
    """Constant costs with non-adaptive penalties keep L = C on every path: S = P * C everywhere,
    which fails if a start pixel is missed or a corner is walked twice."""
    cost = np.full((6, 9, 5), 4, np.uint16)
    guide = np.zeros((6, 9), np.uint8)
    p = O.SgmParams(num_disparities=5, p1=3, p2=10, adaptive_p2=False, eight_paths=True)
    np.testing.assert_array_equal(O.aggregate(cost, guide, p), cost * 8)


def test_bounds_and_thread_independence():
    rng = np.random.default_rng(9)
    cost = rng.integers(0, 64, (23, 31, 16), dtype=np.uint16)
    guide = rng.integers(0, 256, (23, 31), dtype=np.uint8)
    p = O.SgmParams(num_disparities=16, p1=10, p2=120, eight_paths=True)
    S1 = O.aggregate(cost, guide, p, threads=1)
    S8 = O.aggregate(cost, guide, p, threads=8)
    np.testing.assert_array_equal(S1, S8)
    # Per direction 0 <= L - C <= p2max, so C*P <= S <= (C + p2max)*P.
    assert (S1 >= cost * 8).all() and (S1 <= (cost.astype(int) + 120) * 8).all()


def test_path_costs_wrap_at_16_bits_with_large_penalties():
    """stereo.cpp:79-83 casts each path cost and each sum to u16: with a P1 above every path cost and
    adaptive P2 (step penalty P1 + 1 on a flat guide), a disparity whose cost is 63 at every pixel of a
    row keeps its own previous cost (best = L'(k)) while the line minimum stays 0, so along the row
    L_n(k) = 63 (n + 1) mod 2^16, which passes 2^16 after 1040 steps; the four-path sum wraps the same way."""
    w, nd = 1200, 4
    cost = np.zeros((1, w, nd), np.uint16)
    cost[:, :, 1] = 63
    guide = np.full((1, w), 100, np.uint8)
    p = O.SgmParams(num_disparities=nd, census_radius_x=4, census_radius_y=3, p1=100000, p2=120, adaptive_p2=True,
                    eight_paths=False, slant_steps=True)
    n = np.arange(w, dtype=np.int64)
    lr = O.aggregate_direction(cost, guide, p, 1, 0)
    assert np.array_equal(lr[0, :, 1].astype(np.int64), (63 * (n + 1)) % 65536)
    assert (lr[0, :, [0, 2, 3]] == 0).all()
    rl = O.aggregate_direction(cost, guide, p, -1, 0)
    assert np.array_equal(rl[0, :, 1].astype(np.int64), (63 * (w - n)) % 65536)
    s = O.aggregate(cost, guide, p)
    assert np.array_equal(s[0, :, 1].astype(np.int64), (63 * (n + 1) + 63 * (w - n) + 2 * 63) % 65536)
    assert int(lr[0, 1040, 1]) == 47  # 63 * 1041 - 65536
