"""The band-sweep schedule of csrc/sweep_kernels.hip, emulated on the CPU with the kernel's own index
arithmetic: bands of rows launched in order, blocks of `tile` columns, each row's halo of the left
kind on [x0 - h, x0) and of the right kind on [x1, x1 + h) with h = rows left in the band, LDS slots
(sweep_slot) in two row-parity buffers, the band's first predecessors loaded from the seeds into
state buffer 1, each row's costs and penalty words staged in LDS by the block (costs over the row's
columns, penalty words over them and, for the predecessor row, one column more on each side), and
seed rows in two band-parity buffers.

Checked for every access: LDS slots, staged columns and seed, cost, penalty and record offsets lie
inside their buffers; every LDS read finds a slot written on the previous row of the same block (or
preloaded from the seeds), every staged read a column staged for that row, and every seed read a
seed the previous band wrote. The swept sums must equal the oracle's per-direction path costs
summed over the sweep's three directions (stereo.cpp:385-409 via orc_aggregate_direction), for the
real disparities, on tiles and bands that do not divide the image, with padded disparity lanes and
negative minimum disparities. The recurrence here is a plain restatement of stereo.cpp:62-81; the
point of the test is the schedule, which must hand every pixel the predecessor the line walk would.
"""
import numpy as np
import pytest

import oracle as O

SENT = 0x7FFF


def sweep_dir(up, kind):
    return (3, 6, 5)[kind] if up else (2, 4, 7)[kind]


def sweep_slot(kind, x, x0, tile, band):
    return x - x0 if kind == 0 else (tile + (x - x0 + band) if kind == 1 else 2 * tile + band + (x - x0))


def stage_loads(dpl):
    return 4 if dpl <= 8 else (6 if dpl <= 16 else 8)


def sweep_shape_ok(dp, tile, band):
    nc = tile + 2 * band
    return 16 <= tile <= 64 and tile % 16 == 0 and nc + 1 <= 256 and nc * dp <= stage_loads(dp // 16) * 256 * 16


def step(prev, cost, pm, p1, p2, slant, nd):
    """One step of the recurrence on a dp vector whose entries >= nd are pads (0x7FFF)."""
    dp = prev.shape[0]
    pv = prev.astype(np.int64)
    big = np.full(dp + 4, 1 << 40, np.int64)
    big[2:2 + dp] = pv
    nb = np.minimum(big[1:1 + dp], big[3:3 + dp])
    if slant:
        nb = np.minimum(nb, np.minimum(big[0:dp], big[4:4 + dp]))
    best = np.minimum(np.minimum(pv, nb + p1), pm + p2)
    v = cost.astype(np.int64) + best - pm
    v[nd:] = SENT
    return v


def penalty_words(left, p):
    """u32 [y][x]: byte k = step penalty into (x, y) from (x, y) - r_k, r = (1,0), (0,1), (1,1), (1,-1)."""
    h, w = left.shape
    words = np.zeros((h, w), np.uint32)
    img = left.astype(np.int64)
    for k, (rx, ry) in enumerate(((1, 0), (0, 1), (1, 1), (1, -1))):
        for y in range(h):
            for x in range(w):
                qx, qy = x - rx, y - ry
                pen = p.p2
                if p.adaptive_p2 and 0 <= qx < w and 0 <= qy < h:
                    pen = max(p.p1 + 1, p.p2 // (abs(int(img[y, x]) - int(img[qy, qx])) // 8 + 1))
                words[y, x] |= np.uint32(min(pen, 255)) << np.uint32(8 * k)
    return words


def emulate(cost, pen, p, nd, dp, up, tile, band):
    """The swept u16 sums [y][x][dp] of one sweep, with every access checked."""
    assert sweep_shape_ok(dp, tile, band)
    h, w = cost.shape[:2]
    C = np.full((h, w, dp), 0, np.int64)
    C[:, :, :nd] = cost
    S = np.zeros((h, w, dp), np.int64)
    seed = np.zeros((2, 3, w, dp), np.int64)
    seed_ok = np.zeros((2, 3, w), bool)
    nbands = -(-h // band)
    for b in range(nbands):
        t0 = b * band
        nrows = min(band, h - t0)
        seed_ok[b & 1] = False  # this band rewrites its parity
        for ti in range(-(-w // tile)):
            x0, x1 = ti * tile, min(w, ti * tile + tile)
            xb = x0 - band
            ns, ncol = 3 * tile + 2 * band, tile + 2 * band
            lds = np.zeros((2, ns, dp), np.int64)
            lmin = np.zeros((2, ns), np.int64)
            ok = np.zeros((2, ns), bool)
            cst = np.zeros((2, ncol, dp), np.int64)
            cst_ok = np.zeros((2, ncol), bool)
            pe = np.zeros((2, 2 * (ncol + 1)), np.int64)
            pe_ok = np.zeros((2, 2 * (ncol + 1)), bool)

            def row_y(j):
                return h - 1 - (t0 + j) if up else t0 + j

            def row_lo(j):
                return max(0, x0 - (nrows - 1 - j))

            def row_hi(j):
                return min(w, x1 + (nrows - 1 - j))

            def stage(j):
                y = row_y(j)
                yp = y + 1 if up else y - 1
                lo, hi = row_lo(j), row_hi(j)
                par = j & 1
                cst_ok[par] = False
                pe_ok[par] = False
                n16 = (hi - lo) * dp // 16
                assert n16 <= stage_loads(dp // 16) * 256
                for x in range(lo, hi):
                    k = x - xb
                    assert 0 <= k < ncol
                    cst[par, k] = C[y, x]
                    cst_ok[par, k] = True
                    pe[par, k] = int(pen[y, x])
                    pe_ok[par, k] = True
                assert hi - lo <= 256
                if 0 <= yp < h:
                    qs = range(max(0, lo - 1), min(w, hi + 1))
                    assert len(qs) <= 256
                    for xq in qs:
                        k = ncol + 1 + (xq - xb)
                        assert ncol + 1 <= k < 2 * (ncol + 1)
                        pe[par, k] = int(pen[yp, xq])
                        pe_ok[par, k] = True

            # Band start: row 0's predecessors from the seeds into state buffer 1, row 0 staged.
            if t0 > 0:
                h0 = nrows - 1
                lo1, hi2 = max(0, x0 - h0 - 1), min(w, x1 + h0 + 1)
                items = [(0, x) for x in range(x0, x1)] + [(1, x) for x in range(lo1, x1 - 1)] + \
                        [(2, x) for x in range(x0 + 1, hi2)]
                for kind, x in items:
                    par = (b + 1) & 1
                    assert seed_ok[par, kind, x], ("seed not written", b, kind, x)
                    v = seed[par, kind, x].copy()
                    v[nd:] = SENT
                    sl = sweep_slot(kind, x, x0, tile, band)
                    assert 0 <= sl < ns
                    lds[1, sl] = v & 0xFF
                    lmin[1, sl] = int(v.min())
                    ok[1, sl] = True
            stage(0)
            for j in range(nrows):
                ts = t0 + j
                y = row_y(j)
                lo, hi = row_lo(j), row_hi(j)
                ok[j & 1] = False

                def item(kind, x):
                    xp = x + (-1 if kind == 1 else (1 if kind == 2 else 0))
                    sw = sweep_slot(kind, x, x0, tile, band)
                    lo_s = (0, tile, 2 * tile + band)[kind]
                    hi_s = (tile, 2 * tile + band, ns)[kind]
                    assert lo_s <= sw < hi_s, (kind, x, sw)
                    assert 0 <= x < w and 0 <= y < h
                    assert cst_ok[j & 1, x - xb], ("cost not staged", j, x)
                    cc = cst[j & 1, x - xb]
                    if ts == 0 or xp < 0 or xp >= w:
                        v = cc.copy()
                        v[nd:] = SENT
                    else:
                        sr = sweep_slot(kind, xp, x0, tile, band)
                        assert lo_s <= sr < hi_s and ok[(j - 1) & 1, sr], ("LDS slot not written", j, kind, xp)
                        prev = lds[(j - 1) & 1, sr].copy()
                        prev[nd:] = SENT
                        pm = int(lmin[(j - 1) & 1, sr])
                        d = sweep_dir(up, kind)
                        axis = d >> 1
                        k = (x - xb) if (d & 1) == 0 else ncol + 1 + (xp - xb)
                        assert pe_ok[j & 1, k], ("penalty not staged", j, kind, x)
                        p2 = (int(pe[j & 1, k]) >> (8 * axis)) & 0xFF
                        v = step(prev, cc, pm, p.p1, p2, p.slant_steps, nd)
                    assert (v[:nd] <= 255).all() and (v[:nd] >= 0).all()
                    lds[j & 1, sw] = v & 0xFF
                    lmin[j & 1, sw] = int(v.min())
                    ok[j & 1, sw] = True
                    if j == nrows - 1 and x0 <= x < x1:
                        seed[b & 1, kind, x] = v & 0xFF
                        seed_ok[b & 1, kind, x] = True
                    return v

                for x in range(x0, x1):
                    s = item(0, x) + item(1, x) + item(2, x)
                    S[y, x] = s & 0xFFFF
                for x in range(lo, x0):
                    item(1, x)
                for x in range(x1, hi):
                    item(2, x)
                if j + 1 < nrows:
                    stage(j + 1)
    return S


CASES = [  # w, h, nd, dp, min_d, tile, band, slant, adaptive
    (37, 23, 16, 32, 0, 16, 5, True, True),
    (50, 19, 20, 32, -3, 32, 4, False, True),
    (33, 40, 32, 32, 5, 16, 7, True, False),
    (70, 9, 10, 32, 2, 64, 16, True, True),
    (21, 30, 24, 32, -6, 48, 1, False, False),
    # degenerate shapes with 2-row bands and 16-column tiles
    (1, 1, 16, 32, 0, 16, 2, True, True),
    (1, 40, 32, 32, 0, 16, 2, True, True),
    (40, 1, 16, 32, 0, 16, 2, True, True),
    (17, 3, 16, 32, 1, 16, 2, True, True),
    (9, 33, 16, 32, -2, 16, 2, False, True),
]


@pytest.mark.parametrize("w,h,nd,dp,min_d,tile,band,slant,adaptive", CASES)
def test_band_sweeps_reproduce_the_line_walk(w, h, nd, dp, min_d, tile, band, slant, adaptive):
    rng = np.random.default_rng(w * 1000 + h)
    left = rng.integers(0, 256, (h, w), dtype=np.uint8)
    right = rng.integers(0, 256, (h, w), dtype=np.uint8)
    p = O.SgmParams(min_disparity=min_d, num_disparities=nd, census_radius_x=2, census_radius_y=1, p1=10, p2=120,
                    eight_paths=True, adaptive_p2=adaptive, slant_steps=slant)
    cl, cr = O.census(left, 2, 1), O.census(right, 2, 1)
    cost = O.cost_volume(cl, cr, min_d, nd, 15)
    pen = penalty_words(left, p)
    for up, dirs in ((False, ((0, 1), (1, 1), (-1, 1))), (True, ((0, -1), (1, -1), (-1, -1)))):
        want = np.zeros((h, w, nd), np.uint16)
        for dx, dy in dirs:
            want = O.aggregate_direction(cost, left, p, dx, dy, sum_in=want)
        got = emulate(cost, pen, p, nd, dp, up, tile, band)
        np.testing.assert_array_equal(got[:, :, :nd], want.astype(np.int64), err_msg=f"up={up}")
