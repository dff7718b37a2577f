"""Rule-by-rule pins of the oracle's refinement stages (libs/depth/src/stereo.cpp:151-207, 243-284).

A plain Python statement of refine_level and refine_detail, written from the reference's source and
built only on oracle functions pinned elsewhere (zncc_slanted / zncc_window against closed forms and
numpy's correlation, median3 against a brute-force sort, subpixel_offset against the parabola vertex),
must equal the C oracle bit for bit on crafted inputs. The inputs are chosen so that each plausible slip
of a rule changes the output: the test plants every slip in the Python statement and checks that the
result moves, so a C oracle carrying that slip could not pass. The slips:

* the candidate count nd = min(127, 2 * search / 0.5 + 1) and its first candidate 2 d0 - search;
* the slope clamp to [-0.9, 0.9] and the gate |neighbour - centre| < 2 of the coarse slopes;
* the kStep factor on the subpixel offset of refine_level;
* refine_detail's candidate count K = min(16, lround(search / step)), its gradient gate (< 1) and
  clamp, and its keep rule (best_k <= 0, best_k + 1 >= nc, or a neighbour score of -2 keeps the
  first estimate).
"""
import numpy as np
from scipy import ndimage

import oracle as O

F = np.float32
INVALID = F(-1.0e9)


def valid(d):
    return d > F(-1.0e8)


def textured(h, w, seed, sigma=1.0):
    rng = np.random.default_rng(seed)
    img = ndimage.gaussian_filter(rng.normal(128, 60, (h, w)), sigma)
    return np.clip(img, 0, 255).astype(np.uint8)


def shifted_right(left, disp):
    """A right image whose row y at x holds left at x + disp(x, y) (bilinear), so left pixel x matches
    right pixel x - disp: a scene with the disparity map `disp`."""
    h, w = left.shape
    out = np.zeros_like(left)
    for y in range(h):
        xs = np.arange(w, dtype=np.float64)
        src = xs + disp[y]
        out[y] = np.clip(np.interp(src, xs, left[y].astype(np.float64)), 0, 255).round().astype(np.uint8)
    return out


def refine_level_py(left, right, coarse_raw, sp, subpixel, *, clamp=F(0.9), gate=F(2.0), kstep=F(0.5), nd_of=None,
                    lo_of=None, subpixel_factor=True):
    """refine_level, stereo.cpp:151-207, one pixel at a time in float32 as the reference computes."""
    h, w = left.shape
    coarse = O.median3(coarse_raw)
    ch, cw = coarse.shape
    r, search = sp.refine_zncc_radius, sp.refine_search_radius
    out_d = np.full((h, w), INVALID, F)
    out_c = np.zeros((h, w), F)
    nd = min(127, int(F(2 * search) / kstep) + 1) if nd_of is None else nd_of(search)

    def slope(cx, cy):
        gx = gy = F(0.0)
        c = coarse[cy, cx]
        if cx > 0 and cx + 1 < cw and valid(coarse[cy, cx - 1]) and valid(coarse[cy, cx + 1]):
            g = F(0.5) * (coarse[cy, cx + 1] - coarse[cy, cx - 1])
            if abs(coarse[cy, cx + 1] - c) < gate and abs(c - coarse[cy, cx - 1]) < gate:
                gx = g
        if cy > 0 and cy + 1 < ch and valid(coarse[cy - 1, cx]) and valid(coarse[cy + 1, cx]):
            g = F(0.5) * (coarse[cy + 1, cx] - coarse[cy - 1, cx])
            if abs(coarse[cy + 1, cx] - c) < gate and abs(c - coarse[cy - 1, cx]) < gate:
                gy = g
        return F(min(max(gx, -clamp), clamp)), F(min(max(gy, -clamp), clamp))

    for y in range(r, h - r):
        cy = min(y // 2, ch - 1)
        for x in range(r, w - r):
            cx = min(x // 2, cw - 1)
            d0 = coarse[cy, cx]
            if not valid(d0):
                continue
            gx, gy = slope(cx, cy)
            center = F(2.0) * d0
            lo = F(center - F(search)) if lo_of is None else lo_of(center, search)
            scores = [F(O.zncc_slanted(left, right, x, y, float(F(lo + kstep * F(k))), float(gx), float(gy), r))
                      for k in range(nd)]
            best, best_k = F(-2.0), -1
            for k, s in enumerate(scores):
                if s > best:
                    best, best_k = s, k
            if best_k < 0 or best < F(sp.refine_min_zncc):
                continue
            if sp.refine_reject_edge_winners and (best_k == 0 or best_k == nd - 1):
                continue
            d = F(lo + kstep * F(best_k))
            if subpixel and 0 < best_k < nd - 1 and scores[best_k - 1] > -2 and scores[best_k + 1] > -2:
                off = F(O.subpixel_offset(float(F(1) - scores[best_k - 1]), float(F(1) - best),
                                          float(F(1) - scores[best_k + 1])))
                d = F(d + (kstep * off if subpixel_factor else off))
            out_d[y, x] = d
            out_c[y, x] = best
    return out_d, out_c


def refine_detail_py(fl, fr, disparity, sp, subpixel, *, k_of=None, gate=F(1.0), clamp=F(0.9), keep_edges=False):
    """refine_detail, stereo.cpp:243-284, in float32 as the reference computes."""
    first = np.array(disparity, F)
    out = first.copy()
    h, w = first.shape
    H, W = fl.shape
    step, radius = F(sp.detail_step), sp.detail_radius
    K = min(16, int(np.round(np.float64(F(sp.detail_search) / step)))) if k_of is None else k_of(sp)
    nc, n = 2 * K + 1, 2 * radius

    def grad(l, c, r):
        if not valid(l) or not valid(r) or abs(r - c) >= gate or abs(c - l) >= gate:
            return F(0.0)
        return F(min(max(F(0.5) * (r - l), -clamp), clamp))

    for y in range(1, h - 1):
        y0 = 2 * y - radius + 1
        if y0 < 0 or y0 + n > H:
            continue
        for x in range(1, w - 1):
            d = first[y, x]
            if not valid(d):
                continue
            x0 = 2 * x - radius + 1
            if x0 < 0 or x0 + n > W:
                continue
            gx = grad(first[y, x - 1], d, first[y, x + 1])
            gy = grad(first[y - 1, x], d, first[y + 1, x])
            cx, cy = F(2.0) * F(x) + F(0.5), F(2.0) * F(y) + F(0.5)
            D0 = F(2.0) * d
            scores = [F(O.zncc_window(fl, fr, x0, y0, n, float(cx), float(cy), float(F(D0 + step * F(k - K))),
                                      float(gx), float(gy))) for k in range(nc)]
            best, best_k = F(-2.0), -1
            for k, s in enumerate(scores):
                if s > best:
                    best, best_k = s, k
            edge = best_k <= 0 or best_k + 1 >= nc
            if (edge and not keep_edges) or best_k < 0 or best < F(sp.detail_min_zncc):
                continue
            if not edge and (scores[best_k - 1] <= -2 or scores[best_k + 1] <= -2):
                continue
            if edge:  # the slip: an edge winner replaces the estimate (without a subpixel offset)
                out[y, x] = F(F(0.5) * F(D0 + step * F(best_k - K)))
                continue
            D = F(D0 + step * F(best_k - K))
            if subpixel:
                D = F(D + step * F(O.subpixel_offset(float(F(1) - scores[best_k - 1]), float(F(1) - best),
                                                     float(F(1) - scores[best_k + 1]))))
            out[y, x] = F(F(0.5) * D)
    return out


def same(a, b):
    return np.array_equal(np.asarray(a, F).view(np.uint32), np.asarray(b, F).view(np.uint32))


def level_case(search):
    """A 40x28 textured pair over a coarse 20x14 map of ramps whose x gradients take values on both
    sides of the clamp (0.9) and of the gate (2, met exactly by the slope-2 ramps, whose 3x3 median keeps
    them), y gradients in bands, and the scene's true disparity near or past the last candidate."""
    rng = np.random.default_rng(100 + search)
    ch, cw = 14, 20
    yy, xx = np.mgrid[0:ch, 0:cw].astype(np.float64)
    slopes = np.array([0.3, 0.95, 2.0, 1.5, 2.6])
    gx = slopes[(xx // 4).astype(int) % len(slopes)] * np.where((yy // 5) % 2 == 0, 1, -1)
    coarse = 8.0 + gx * (xx % 4) + 0.8 * (yy % 5)
    coarse[10:, :] += rng.uniform(-0.2, 0.2, (ch - 10, cw))  # bands of unequal neighbours too
    coarse = coarse.astype(F)
    coarse[2, 5] = INVALID
    left = textured(2 * ch, 2 * cw, 7 + search)
    fine = np.kron(coarse.astype(np.float64), np.ones((2, 2))) * 2.0
    fine[np.kron(coarse, np.ones((2, 2))) < -1e8] = 16.0
    fine += search + 0.25  # near or past the last candidate 2 d0 + search
    right = shifted_right(left, fine)
    return left, right, coarse


def test_refine_level_rules():
    for search in (1, 2, 3, 4):
        left, right, coarse = level_case(search)
        for reject, subpixel in ((False, True), (True, True), (False, False)):
            sp = O.StereoParams(refine_search_radius=search, refine_zncc_radius=2, refine_min_zncc=-1.0,
                                refine_reject_edge_winners=reject)
            ref_d, ref_c = O.refine_level(left, right, coarse, sp, subpixel=subpixel)
            py_d, py_c = refine_level_py(left, right, coarse, sp, subpixel)
            assert (py_d > -1e8).sum() > 100
            assert same(ref_d, py_d) and same(ref_c, py_c), (search, reject, subpixel)
    # Every slip moves the result on these inputs.
    left, right, coarse = level_case(2)
    sp = O.StereoParams(refine_search_radius=2, refine_zncc_radius=2, refine_min_zncc=-1.0,
                        refine_reject_edge_winners=True)
    good = refine_level_py(left, right, coarse, sp, True)[0]
    slips = {
        "one candidate fewer": dict(nd_of=lambda s: int(2 * s / 0.5)),
        "one candidate more": dict(nd_of=lambda s: int(2 * s / 0.5) + 2),
        "first candidate at 2 d0": dict(lo_of=lambda c, s: F(c)),
        "slope clamp 1.0": dict(clamp=F(1.0)),
        "slope clamp 0.8": dict(clamp=F(0.8)),
        "slope gate <= 2": dict(gate=np.nextafter(F(2.0), F(3.0))),
        "slope gate < 1.5": dict(gate=F(1.5)),
        "no kStep on the subpixel offset": dict(subpixel_factor=False),
    }
    for name, kw in slips.items():
        assert not same(good, refine_level_py(left, right, coarse, sp, True, **kw)[0]), name


def detail_case():
    """A 2x full-resolution pair (64x32 -> 32x16 half-resolution estimate) whose estimate has
    gradients on both sides of the gate (1, met exactly by the slope-1 ramps) and of the clamp (0.9), and
    true disparities inside, at and beyond the search range."""
    rng = np.random.default_rng(5)
    h, w = 16, 32
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    g = np.array([0.2, 0.6, 0.95, 1.2, 1.0])[(xx // 5).astype(int) % 5]
    first = 6.0 + g * (xx % 5) + 0.25 * (yy % 4)  # exact ramps: neighbour steps of exactly 1.0 too
    first[8:, :] += 0.3 * np.sin(yy[8:]) + rng.uniform(-0.05, 0.05, (h - 8, w))
    first = first.astype(F)
    first[4, 9] = INVALID
    fl = textured(2 * h, 2 * w, 11)
    true = np.kron(first.astype(np.float64), np.ones((2, 2))) * 2.0
    true[np.kron(first, np.ones((2, 2))) < -1e8] = 12.0
    true += np.where(np.kron(xx, np.ones((2, 2))) % 5 < 2, 0.3, 1.1)  # inside and beyond the +-1 px search
    # The slope-1 ramp (columns 20..24, rows 0..7) sees a scene flat along x whose disparity the search
    # of the ramp's middle pixel (first = 8 + y steps) contains: there the winner is interior and its
    # score depends on whether the unit steps pass the gradient gate.
    for y in range(8):
        true[2 * y:2 * y + 2, 40:50] = 16.3 + 0.5 * (y % 4)
    fr = shifted_right(fl, true)
    return fl, fr, first


def test_refine_detail_rules():
    fl, fr, first = detail_case()
    for search, step, subpixel in ((1.0, 0.5, True), (1.0, 0.25, True), (2.0, 0.5, False), (0.75, 0.25, True)):
        sp = O.StereoParams(detail_enabled=True, detail_radius=3, detail_search=search, detail_step=step,
                            detail_min_zncc=-1.0)
        ref = O.refine_detail(fl, fr, first, sp, subpixel=subpixel)
        py = refine_detail_py(fl, fr, first, sp, subpixel)
        assert (py != first).sum() > 50
        assert same(ref, py), (search, step, subpixel)
    sp = O.StereoParams(detail_enabled=True, detail_radius=3, detail_search=1.0, detail_step=0.5, detail_min_zncc=-1.0)
    good = refine_detail_py(fl, fr, first, sp, True)
    slips = {
        "K one smaller": dict(k_of=lambda p: int(np.round(p.detail_search / p.detail_step)) - 1),
        "K one larger": dict(k_of=lambda p: int(np.round(p.detail_search / p.detail_step)) + 1),
        "edge winners replace the estimate": dict(keep_edges=True),
        "gradient gate <= 1": dict(gate=np.nextafter(F(1.0), F(2.0))),
        "gradient gate < 2": dict(gate=F(2.0)),
        "gradient clamp 1.0": dict(clamp=F(1.0)),
    }
    for name, kw in slips.items():
        assert not same(good, refine_detail_py(fl, fr, first, sp, True, **kw)), name
