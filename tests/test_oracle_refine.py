"""Pins of the oracle's refinement and post-filter stages (libs/depth/src/stereo.cpp:98-326, 465-536):
ZNCC closed forms and numpy's Pearson correlation, recovery of known shifts and slanted planes,
the rejection rules, brute-force filters, connected components against scipy.ndimage.label, and the
reference's own compute_disparity quality bar (tests/depth/test_stereo.cpp:95-119)."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from open_source_einstar_amd.synth import speckle_pair

INVALID = np.float32(-1.0e9)


def textured(h, w, seed, sigma=1.2):
    rng = np.random.default_rng(seed)
    img = ndimage.gaussian_filter(rng.normal(128, 60, (h, w)), sigma)
    return np.clip(img, 0, 255).astype(np.uint8)


# ---------------------------------------------------------------- ZNCC

def test_zncc_closed_forms():
    a = textured(40, 60, 1)
    assert O.zncc_slanted(a, a, 30, 20, 0.0, 0.0, 0.0, 3) == pytest.approx(1.0, abs=1e-6)
    b = (a // 2 + 7).astype(np.uint8)  # affine in intensity (up to the integer division)
    assert O.zncc_slanted(a, b, 30, 20, 0.0, 0.0, 0.0, 3) > 0.99
    inv = (255 - a).astype(np.uint8)
    assert O.zncc_slanted(a, inv, 30, 20, 0.0, 0.0, 0.0, 3) == pytest.approx(-1.0, abs=1e-6)
    flat = np.full_like(a, 90)
    assert O.zncc_slanted(flat, a, 30, 20, 0.0, 0.0, 0.0, 3) == -1.0
    assert O.zncc_slanted(a, flat, 30, 20, 0.0, 0.0, 0.0, 3) == -1.0
    # The window leaves b: x0 < 0 on the left, x0 + 1 >= width on the right.
    assert O.zncc_slanted(a, a, 5, 20, 3.0, 0.0, 0.0, 3) == -2.0
    assert O.zncc_slanted(a, a, 55, 20, -2.0, 0.0, 0.0, 3) == -2.0
    assert O.zncc_window(a, a, 10, 10, 10, 15.5, 15.5, 0.0, 0.0, 0.0) == pytest.approx(1.0, abs=1e-6)
    assert O.zncc_window(a, a, 0, 10, 10, 5.5, 15.5, 0.5, 0.0, 0.0) == -2.0


def bilinear_row(img, y, xb):
    x0 = np.floor(xb).astype(int)
    f = (xb - x0).astype(np.float32)
    row = img[y].astype(np.float32)
    return row[x0] * (np.float32(1) - f) + row[x0 + 1] * f


def test_zncc_is_pearson_correlation_of_the_slanted_samples():
    rng = np.random.default_rng(2)
    a, b = textured(50, 80, 3), textured(50, 80, 4)
    for _ in range(20):
        x, y, r = int(rng.integers(20, 60)), int(rng.integers(8, 42)), int(rng.integers(1, 5))
        d, gx, gy = float(rng.uniform(-4, 4)), float(rng.uniform(-0.9, 0.9)), float(rng.uniform(-0.9, 0.9))
        sa, sb = [], []
        for v in range(-r, r + 1):
            u = np.arange(-r, r + 1)
            xb = (x + u).astype(np.float32) - (np.float32(d) + np.float32(gx) * u.astype(np.float32)
                                               + np.float32(gy) * np.float32(v))
            sb.append(bilinear_row(b, y + v, xb))
            sa.append(a[y + v, x - r:x + r + 1].astype(np.float64))
        want = np.corrcoef(np.concatenate(sa), np.concatenate(sb).astype(np.float64))[0, 1]
        assert O.zncc_slanted(a, b, x, y, d, gx, gy, r) == pytest.approx(want, abs=2e-6)


def test_zncc_window_is_pearson_correlation():
    rng = np.random.default_rng(5)
    a, b = textured(60, 90, 6), textured(60, 90, 7)
    for _ in range(20):
        n = int(rng.integers(2, 12))
        x0, y0 = int(rng.integers(10, 70)), int(rng.integers(0, 60 - n))
        cx, cy = x0 + n / 2 - 0.5, y0 + n / 2 - 0.5
        D, gx, gy = float(rng.uniform(-5, 5)), float(rng.uniform(-0.9, 0.9)), float(rng.uniform(-0.9, 0.9))
        sa, sb = [], []
        for j in range(n):
            Y = y0 + j
            X = np.arange(x0, x0 + n)
            dy = np.float32(gy) * (np.float32(Y) - np.float32(cy))
            xb = X.astype(np.float32) - (np.float32(D) + np.float32(gx) * (X.astype(np.float32) - np.float32(cx)) + dy)
            sb.append(bilinear_row(b, Y, xb))
            sa.append(a[Y, x0:x0 + n].astype(np.float64))
        want = np.corrcoef(np.concatenate(sa), np.concatenate(sb).astype(np.float64))[0, 1]
        got = O.zncc_window(a, b, x0, y0, n, np.float32(cx), np.float32(cy), D, gx, gy)
        assert got == pytest.approx(want, abs=2e-6)


# ---------------------------------------------------------------- refine_level

def shifted_pair(h, w, shift, seed):
    """left(x) = right(x - shift) exactly, for an integer shift."""
    big = textured(h, w + shift, seed)
    return big[:, :w].copy(), big[:, shift:shift + w].copy()


def test_refine_level_recovers_an_integer_shift_from_a_coarse_estimate():
    left, right = shifted_pair(64, 96, 6, 10)
    coarse = np.full((32, 48), 3.0, np.float32)  # 6 / 2
    d, c = O.refine_level(left, right, coarse, O.StereoParams())
    interior = d[3:-3, 20:-3]
    assert (np.abs(interior - 6.0) < 0.1).mean() > 0.97  # parabola bias of a 0.5 px search step
    assert (c[3:-3, 20:-3] > 0.95).mean() > 0.95
    # Borders of the window and the pixels whose windows leave the right image stay invalid.
    assert (d[:3] == INVALID).all() and (d[-3:] == INVALID).all() and (d[:, :3] == INVALID).all()


def test_refine_level_follows_a_slanted_plane():
    """A plane d = 4 + 0.3 x' on a smooth texture: the coarse gradient slants the window."""
    h, w = 60, 120
    tex = textured(h, w + 60, 11, sigma=2.0).astype(np.float64)
    xs = np.arange(w, dtype=np.float64)
    d_true = 4.0 + 0.3 * (xs - 40) / 2 + 10  # positive everywhere
    right = tex[:, :w].astype(np.uint8)
    left = np.empty((h, w), np.uint8)
    for x in range(w):  # left(x) = right(x - d(x)), linear interpolation of the texture
        xr = x - d_true[x]
        x0 = int(np.floor(xr))
        f = xr - x0
        left[:, x] = np.clip(np.round(tex[:, x0] * (1 - f) + tex[:, x0 + 1] * f), 0, 255)
    coarse = np.tile((d_true[::2] / 2).astype(np.float32)[None, :], (h // 2, 1))
    d, _ = O.refine_level(left, right, coarse, O.StereoParams())
    ok = d[5:-5, 40:-5] > -1e8
    err = np.abs(d[5:-5, 40:-5] - d_true[None, 40:-5])[ok]
    assert ok.mean() > 0.9 and np.median(err) < 0.1


def test_refine_level_rejection_rules():
    left, right = shifted_pair(40, 80, 6, 12)
    off = np.full((20, 40), 3.0 + 1.5, np.float32)  # centre 9: the true 6 lies outside [7, 11]
    d, _ = O.refine_level(left, right, off, O.StereoParams())
    assert (d[3:-3, 20:-3] == INVALID).mean() > 0.9  # best at the search edge: rejected
    d2, c2 = O.refine_level(left, right, off, O.StereoParams(refine_reject_edge_winners=False))
    sel = d2[3:-3, 20:-3]
    assert (np.abs(sel - 7.0) < 0.01).mean() > 0.9  # kept at the edge, no subpixel step there
    d3, _ = O.refine_level(left, right, np.full((20, 40), 3.0, np.float32), O.StereoParams(refine_min_zncc=1.5))
    assert (d3 == INVALID).all()


def test_refine_level_median_filters_the_coarse_input():
    left, right = shifted_pair(40, 80, 6, 13)
    coarse = np.full((20, 40), 3.0, np.float32)
    coarse[10, 20] = 9.0  # an outlier the 3x3 median removes
    d_out, _ = O.refine_level(left, right, coarse, O.StereoParams())
    d_ref, _ = O.refine_level(left, right, np.full((20, 40), 3.0, np.float32), O.StereoParams())
    np.testing.assert_array_equal(d_out[18:24, 38:44], d_ref[18:24, 38:44])


# ---------------------------------------------------------------- refine_detail

def test_refine_detail_improves_a_half_resolution_estimate():
    p = speckle_pair(320, 256, 64, 21)
    hl, hr = O.downsample2(p.left), O.downsample2(p.right)
    gt_half = p.gt_disparity[::2, ::2] / 2
    noisy = (gt_half + np.random.default_rng(0).uniform(-0.3, 0.3, gt_half.shape)).astype(np.float32)
    out = O.refine_detail(p.left, p.right, noisy, O.StereoParams())
    inner = (slice(4, -4), slice(40, -4))
    e0 = np.abs(noisy - gt_half)[inner]
    e1 = np.abs(out - gt_half)[inner]
    assert np.median(e1) < 0.5 * np.median(e0)
    # Frame borders keep the first estimate.
    np.testing.assert_array_equal(out[0], noisy[0])
    np.testing.assert_array_equal(out[:, 0], noisy[:, 0])
    # No clear peak (a flat right image: every score is -1): everything keeps its value.
    flat = np.full_like(p.right, 77)
    np.testing.assert_array_equal(O.refine_detail(p.left, flat, noisy, O.StereoParams()), noisy)
    # Disabled: unchanged.
    np.testing.assert_array_equal(O.refine_detail(p.left, p.right, noisy, O.StereoParams(detail_enabled=False)), noisy)


# ---------------------------------------------------------------- filters

def test_occlusion_filter_known_answers():
    sp = O.StereoParams(filter_spikes=False)
    row = np.full((1, 12), INVALID, np.float32)
    row[0, 8] = 4.0    # x - d = 4: claims right pixel 4
    row[0, 6] = 2.2    # x - d = 3.8 -> 4: background, 1.8 behind the claimant: occluded
    row[0, 9] = 4.6    # x - d = 4.4 -> 4: nearest claimant of 4 is now 4.6; 4.0 is within 1.0: kept
    row[0, 3] = 0.5    # x - d = 2.5 -> 3 (lround, away from zero); alone: kept
    row[0, 2] = -0.5   # x - d = 2.5 -> 3: 1.0 behind 0.5 is not more than the tolerance: kept
    row[0, 1] = 3.0    # x - d = -2: outside the image: kept
    out = O.filter_disparity(row, sp)
    assert out[0, 6] == INVALID
    for x in (8, 9, 3, 2, 1):
        assert out[0, x] == row[0, x], x
    row[0, 2] = -0.51  # now more than 1.0 behind: occluded
    assert O.filter_disparity(row, sp)[0, 2] == INVALID


def test_spike_filter_against_brute_force():
    rng = np.random.default_rng(3)
    sp = O.StereoParams(filter_occlusion=False)
    for trial in range(5):
        d = rng.normal(10, 1.5, (17, 23)).astype(np.float32)
        d[rng.random(d.shape) < 0.3] = INVALID
        out = O.filter_disparity(d, sp)
        want = d.copy()
        h, w = d.shape
        for y in range(h):
            for x in range(w):
                if d[y, x] <= -1e8:
                    continue
                nb = sorted(float(d[yy, xx]) for yy in range(max(0, y - 1), min(h, y + 2))
                            for xx in range(max(0, x - 1), min(w, x + 2)) if d[yy, xx] > -1e8)
                if len(nb) >= 3 and abs(float(d[y, x]) - nb[len(nb) // 2]) > 1.0:
                    want[y, x] = INVALID
        np.testing.assert_array_equal(out, want, err_msg=f"trial {trial}")


# ---------------------------------------------------------------- speckles

def test_speckle_removal_matches_scipy_labels_on_quantised_maps():
    rng = np.random.default_rng(4)
    for trial in range(4):
        d = rng.integers(0, 3, (40, 50)).astype(np.float32)
        d = ndimage.median_filter(d, 3)  # blobs of equal values
        d[rng.random(d.shape) < 0.1] = INVALID
        sp = O.StereoParams(speckle_max_region_size=12, speckle_max_diff=0.5)  # only equal values connect
        out = O.remove_speckles(d, sp)
        want = d.copy()
        for v in np.unique(d[d > -1e8]):
            lab, n = ndimage.label(d == v)  # 4-connectivity
            sizes = np.bincount(lab.ravel())
            small = (sizes < 12) & (np.arange(len(sizes)) > 0)
            want[small[lab]] = INVALID
        np.testing.assert_array_equal(out, want, err_msg=f"trial {trial}")


def union_find_reference(d, max_diff, max_region):
    h, w = d.shape
    parent = list(range(h * w))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    valid = d > -1e8
    for y in range(h):
        for x in range(w):
            if not valid[y, x]:
                continue
            for yy, xx in ((y, x + 1), (y + 1, x)):
                if yy < h and xx < w and valid[yy, xx] and abs(float(d[yy, xx]) - float(d[y, x])) <= max_diff:
                    a, b = find(y * w + x), find(yy * w + xx)
                    parent[max(a, b)] = min(a, b)
    roots = np.array([find(i) for i in range(h * w)])
    sizes = np.bincount(roots[valid.ravel()], minlength=h * w)
    out = d.copy().ravel()
    out[valid.ravel() & (sizes[roots] < max_region)] = INVALID
    return out.reshape(h, w)


def test_speckle_removal_matches_union_find_on_float_maps():
    rng = np.random.default_rng(8)
    for trial in range(4):
        d = np.cumsum(rng.normal(0, 0.6, (30, 40)), axis=1).astype(np.float32)
        d[rng.random(d.shape) < 0.15] = INVALID
        sp = O.StereoParams(speckle_max_region_size=25, speckle_max_diff=1.0)
        np.testing.assert_array_equal(O.remove_speckles(d, sp), union_find_reference(d, 1.0, 25))


# ---------------------------------------------------------------- compute_disparity

SGM = dict(census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=True)


def test_compute_disparity_quality_on_speckle():
    """The reference's bar (test_stereo.cpp:95-119: > 85% valid, small median error, < 1% outliers),
    in disparity units, on a scene of two or three tilted planes (the reference scores a tilted plane
    and a sphere), two pyramid levels."""
    p = speckle_pair(640, 512, 96, 31, min_cells=2, max_cells=3)
    sg = O.SgmParams(min_disparity=max(0, p.min_disparity // 4 - 1), num_disparities=32, **SGM)
    r = O.compute_disparity(p.left, p.right, sg, O.StereoParams(speckle_max_region_size=50))
    d, c = r["disparity"], r["confidence"]
    gt = p.gt_disparity
    sel = np.zeros_like(gt, bool)
    sel[8:-8, 8:-8] = True
    sel &= (np.arange(gt.shape[1])[None, :] - gt) > 8  # correspondences inside the right image
    valid = d > -1e8
    err = np.abs(d - gt)[sel & valid]
    assert valid[sel].mean() > 0.85
    assert np.median(err) < 0.15
    assert (err > 1.0).mean() < 0.01
    assert ((c > 0) == valid).all() and (c[valid] <= 1.0).all()


def test_compute_disparity_is_the_composition_of_its_stages():
    p = speckle_pair(320, 256, 64, 41)
    fl = np.repeat(np.repeat(p.left, 2, 0), 2, 1)
    fr = np.repeat(np.repeat(p.right, 2, 0), 2, 1)
    sg = O.SgmParams(min_disparity=max(0, p.min_disparity // 4 - 1), num_disparities=24, **SGM)
    sp = O.StereoParams(speckle_max_region_size=30)
    got = O.compute_disparity(p.left, p.right, sg, sp, full_left=fl, full_right=fr)
    l1, r1 = O.downsample2(p.left), O.downsample2(p.right)
    l2, r2 = O.downsample2(l1), O.downsample2(r1)
    d = O.sgm(l2, r2, sg)["disparity"]
    d, _ = O.refine_level(l1, r1, d, sp)
    d, c = O.refine_level(p.left, p.right, d, sp)
    d = O.refine_detail(fl, fr, d, sp)
    d = O.remove_speckles(O.filter_disparity(d, sp), sp)
    c[d <= -1e8] = 0
    np.testing.assert_array_equal(got["disparity"], d)
    np.testing.assert_array_equal(got["confidence"], c)
    # Level 0: the SGM disparity, confidence 1 where valid, then the filters.
    sp0 = O.StereoParams(pyramid_levels=0, detail_enabled=False, filter_occlusion=False, filter_spikes=False,
                         speckle_max_region_size=0)
    sg0 = O.SgmParams(min_disparity=p.min_disparity, num_disparities=64, **SGM)
    r0 = O.compute_disparity(p.left, p.right, sg0, sp0)
    d0 = O.sgm(p.left, p.right, sg0)["disparity"]
    np.testing.assert_array_equal(r0["disparity"], d0)
    np.testing.assert_array_equal(r0["confidence"], (d0 > -1e8).astype(np.float32))
