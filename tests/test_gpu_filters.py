"""The disparity post-filters as building blocks (sgm_remove_speckles_device, sgm_filter_disparity_device)
against the oracle's remove_speckles and filter_disparity (libs/depth/src/stereo.cpp:465-494, 286-326) on
crafted maps: shapes whose regions cross many 64x64 labelling tiles along long chains (serpentines,
spirals, combs joined only across tile borders), a checkerboard of single pixels, one region covering
the whole map, differences on both sides of max_diff, ragged, wide and tall sizes, and region limits from
none to all."""
import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm

pytestmark = pytest.mark.gpu

INVALID = O.INVALID_DISPARITY


def same_bits(a, b, tag):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    if not np.array_equal(a.view(np.uint32), b.view(np.uint32)):
        idx = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
        raise AssertionError(f"{tag}: {len(idx)} differences, first at {idx[0].tolist()}: oracle {a[tuple(idx[0])]} "
                             f"gpu {b[tuple(idx[0])]}")


def serpentine(w, h, pitch):
    """A one-pixel-wide path down the map: full rows every `pitch` rows joined at alternating ends."""
    d = np.full((h, w), INVALID, np.float32)
    for i, y in enumerate(range(0, h, pitch)):
        d[y, :] = 10.0
        if y + pitch < h:
            x = w - 1 if i % 2 == 0 else 0
            d[y:y + pitch + 1, x] = 10.0
    return d


def spiral(w, h):
    d = np.full((h, w), INVALID, np.float32)
    x0, y0, x1, y1 = 0, 0, w - 1, h - 1
    while x0 <= x1 and y0 <= y1:
        d[y0, x0:x1 + 1] = 20.0
        d[y0:y1 + 1, x1] = 20.0
        d[y1, x0:x1 + 1] = 20.0
        if y0 + 2 <= y1:
            d[y0 + 2:y1 + 1, x0] = 20.0
        x0, y0, x1, y1 = x0 + 2, y0 + 2, x1 - 2, y1 - 2
        if x0 <= x1 and y0 <= y1:
            d[y0 - 1, x0 - 2 if x0 >= 2 else x0] = 20.0  # the turn into the next ring
    return d


def comb(w, h):
    """Vertical teeth of one pixel every other column, joined by a top row only at tile borders, so that
    most unions happen across the 64x64 tiles."""
    d = np.full((h, w), INVALID, np.float32)
    d[1:, ::2] = 5.0
    for x in range(63, w - 1, 64):
        d[0, x - 1:x + 2] = 5.0
    return d


def noisy_regions(w, h, seed, cells=(3, 12)):
    """Piecewise-constant regions with sub-threshold noise, invalid holes and isolated pixels."""
    rng = np.random.default_rng(seed)
    cs = rng.integers(cells[0], cells[1] + 1)
    base = rng.uniform(5, 60, ((h + cs - 1) // cs, (w + cs - 1) // cs)).astype(np.float32)
    d = np.kron(base, np.ones((cs, cs), np.float32))[:h, :w]
    d = d + rng.uniform(-0.3, 0.3, (h, w)).astype(np.float32)
    d[rng.random((h, w)) < 0.08] = INVALID
    return np.ascontiguousarray(d, np.float32)


MAPS = {
    "serpentine-300x260": lambda: serpentine(300, 260, 2),
    "serpentine-1000x130": lambda: serpentine(1000, 130, 3),
    "spiral-257x190": lambda: spiral(257, 190),
    "comb-640x100": lambda: comb(640, 100),
    "checkerboard-130x70": lambda: np.where((np.indices((70, 130)).sum(0) % 2) == 0, np.float32(7), INVALID).astype(np.float32),
    "one-region-500x300": lambda: np.full((300, 500), 12.5, np.float32),
    "ramp-at-max-diff-200x90": lambda: (np.arange(200, dtype=np.float32)[None, :] * np.float32(1.0) +
                                        np.zeros((90, 1), np.float32)),
    "noisy-250x170": lambda: noisy_regions(250, 170, 3),
    "noisy-1280x1024": lambda: noisy_regions(1280, 1024, 4, cells=(2, 9)),
    "noisy-5000x20": lambda: noisy_regions(5000, 20, 5),
    "noisy-20x3000": lambda: noisy_regions(20, 3000, 6),
    "single-pixel-1x1": lambda: np.full((1, 1), 3.0, np.float32),
    "row-1x777": lambda: noisy_regions(777, 1, 7, cells=(1, 4)),
    "column-1x3000": lambda: noisy_regions(1, 3000, 8, cells=(1, 40)),
    # a one-pixel staircase down the diagonal: one 4-connected region crossing tiles corner to corner
    "staircase-200x200": lambda: np.where(np.equal(*np.indices((200, 200))) | np.equal(np.indices((200, 200))[0],
                                                                                          np.indices((200, 200))[1] + 1),
                                          np.float32(9), INVALID).astype(np.float32),
    # the diagonal alone: its pixels touch only at corners, so each is a region of one
    "diagonal-150x150": lambda: np.where(np.equal(*np.indices((150, 150))), np.float32(9), INVALID).astype(np.float32),
}


@pytest.mark.parametrize("name", list(MAPS))
@pytest.mark.parametrize("max_region,max_diff", [(0, 1.0), (2, 1.0), (60, 1.0), (200, 0.5), (10**9, 1.0), (300, 0.0)],
                         ids=["none", "singles", "60", "200-diff0.5", "all", "300-diff0"])
def test_remove_speckles_equals_oracle(name, max_region, max_diff):
    d = MAPS[name]()
    print(f"speckles {name} {d.shape} max_region {max_region} max_diff {max_diff}", flush=True)
    ref = O.remove_speckles(d, O.StereoParams(speckle_max_region_size=max_region, speckle_max_diff=max_diff))
    got = sgm.remove_speckles(d, max_region, max_diff)
    same_bits(ref, got, f"{name} max_region {max_region} max_diff {max_diff}")


def test_speckles_unit_difference_ramp_is_one_region_or_none():
    """A ramp that steps by exactly max_diff per column is one region; a hair above, every column its own."""
    d = np.tile(np.arange(0, 96, dtype=np.float32) * np.float32(0.75), (40, 1))
    for md, expect_all_removed in ((0.75, False), (np.nextafter(np.float32(0.75), np.float32(0)), True)):
        ref = O.remove_speckles(d, O.StereoParams(speckle_max_region_size=200, speckle_max_diff=float(md)))
        got = sgm.remove_speckles(d, 200, float(md))
        same_bits(ref, got, f"ramp max_diff {md}")
        assert (got <= -1e8).all() == expect_all_removed


@pytest.mark.parametrize("name", ["noisy-250x170", "noisy-1280x1024", "noisy-5000x20", "checkerboard-130x70", "row-1x777"])
@pytest.mark.parametrize("occ,occ_tol,spk,spk_thr", [(True, 1.0, True, 1.0), (True, 0.25, False, 1.0), (False, 1.0, True, 0.3),
                                                     (False, 1.0, False, 1.0)],
                         ids=["both", "occlusion", "spikes", "neither"])
def test_filter_disparity_equals_oracle(name, occ, occ_tol, spk, spk_thr):
    d = MAPS[name]()
    print(f"filters {name} {d.shape} occlusion {occ}/{occ_tol} spikes {spk}/{spk_thr}", flush=True)
    sp = O.StereoParams(filter_occlusion=occ, filter_occlusion_tolerance=occ_tol, filter_spikes=spk,
                        filter_spike_threshold=spk_thr)
    ref = O.filter_disparity(d, sp)
    got = sgm.filter_disparity(d, occ, occ_tol, spk, spk_thr)
    same_bits(ref, got, f"{name} filters")


def test_refusals():
    """Null maps, empty sizes, a missing workspace where regions are counted and a width beyond the
    limit are refused with SGM_ERR_INVALID_ARGUMENT (5) before anything is enqueued."""
    L = sgm.lib()
    buf = sgm.DeviceBuffer(4 * 64)
    try:
        assert L.sgm_remove_speckles_device(None, buf.ptr, 8, 8, 50, 1.0, buf.ptr, None) == 5
        assert L.sgm_remove_speckles_device(buf.ptr, buf.ptr, 0, 8, 50, 1.0, buf.ptr, None) == 5
        assert L.sgm_remove_speckles_device(buf.ptr, buf.ptr, 8, 8, 50, 1.0, None, None) == 5
        assert b"workspace" in L.sgm_last_error(None)
        assert L.sgm_remove_speckles_device(buf.ptr, buf.ptr, 8, 8, 1, 1.0, None, None) == 0  # copies: no workspace
        assert L.sgm_filter_disparity_device(None, 8, 8, 1, 1.0, 1, 1.0, buf.ptr, None) == 5
        assert L.sgm_filter_disparity_device(buf.ptr, 8, 8, 1, 1.0, 1, 1.0, None, None) == 5
        assert L.sgm_filter_disparity_device(buf.ptr, 40000, 1, 1, 1.0, 0, 1.0, None, None) == 5
        assert L.sgm_speckle_workspace_bytes(100, 30) == 12 * 3000 and L.sgm_speckle_workspace_bytes(0, 5) == 0
    finally:
        buf.free()
