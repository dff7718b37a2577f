"""compute_disparity on the GPU (pyramid, slanted ZNCC refinement per level, detail refinement,
occlusion/spike filters, speckle removal) against the oracle's restatement of
libs/depth/src/stereo.cpp:98-536, stage by stage and composed, from rectified and from raw frames.

Bar: disparity and confidence bit-identical (the scores accumulate in double in the reference's
order); the stated tolerance for float disparities, 1e-4 px with identical validity, is what the
assertion message reports if they are not."""
import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import rectification_maps, speckle_pair

pytestmark = pytest.mark.gpu

TOL = 1e-4
SGM = dict(census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=True)


def assert_maps(ref, got, tag):
    for f in ("disparity", "confidence"):
        a, b = np.asarray(ref[f]), np.asarray(got[f])
        assert a.shape == b.shape, (tag, f)
        if np.array_equal(a.view(np.uint32), b.view(np.uint32)):
            continue
        va, vb = a > -1e8, b > -1e8
        diff = np.abs(a - b)[va & vb] if f == "disparity" else np.abs(a - b)
        raise AssertionError(f"{tag} {f}: {int((a != b).sum())} pixels differ, validity differs at "
                             f"{int((va != vb).sum())}, worst {float(diff.max()) if diff.size else 0:.3g} "
                             f"(tolerance {TOL})")


def sgm_params(pair, levels, nd_top):
    return dict(min_disparity=pair.min_disparity // (1 << levels) - 1, num_disparities=nd_top, **SGM)


def run(w, h, nd, seed, stereo: dict, nd_top, full=False, cells=(2, 4)):
    pair = speckle_pair(w, h, nd, seed, min_cells=cells[0], max_cells=cells[1])
    st = O.StereoParams(**stereo)
    kw = sgm_params(pair, st.pyramid_levels, nd_top)
    fl = fr = None
    if full:  # a full-resolution pair that downsamples exactly to (left, right)
        fl = np.repeat(np.repeat(pair.left, 2, 0), 2, 1)
        fr = np.repeat(np.repeat(pair.right, 2, 0), 2, 1)
    print(f"compute_disparity {w}x{h} {stereo} full={full}", flush=True)
    ref = O.compute_disparity(pair.left, pair.right, O.SgmParams(**kw), st, full_left=fl, full_right=fr)
    with sgm.HipStereo.create_stereo(sgm.SgmParams(**kw), sgm.StereoParams(**stereo), w, h) as ctx:
        got = ctx.compute_disparity(pair.left, pair.right, fl, fr)
        t = ctx.last_timings()
    assert_maps(ref, got, f"{w}x{h} {stereo}")
    return ref, got, t


OFF = dict(detail_enabled=False, filter_occlusion=False, filter_spikes=False, speckle_max_region_size=0)

STAGES = [
    ("refine-1-level", 320, 256, 64, dict(OFF, pyramid_levels=1), 40, False),
    ("refine-2-levels", 320, 256, 64, dict(OFF, pyramid_levels=2), 24, False),
    ("refine-3-levels-r2-s1", 400, 320, 96, dict(OFF, pyramid_levels=3, refine_zncc_radius=2, refine_search_radius=1), 16, False),
    ("refine-keep-edge-winners", 256, 192, 48, dict(OFF, pyramid_levels=1, refine_reject_edge_winners=False), 32, False),
    ("detail-only", 256, 192, 48, dict(OFF, pyramid_levels=0, detail_enabled=True), 48, True),
    ("detail-after-refine", 320, 256, 64, dict(OFF, pyramid_levels=1, detail_enabled=True), 40, True),
    # partial 16x16 refine and detail blocks (and their LDS tiles) at the right and bottom edges
    ("refine-detail-250x170", 250, 170, 64, dict(OFF, pyramid_levels=1, detail_enabled=True), 40, True),
    ("occlusion", 256, 192, 48, dict(OFF, pyramid_levels=0, filter_occlusion=True), 48, False),
    ("spikes", 256, 192, 48, dict(OFF, pyramid_levels=0, filter_spikes=True), 48, False),
    ("speckles", 256, 192, 48, dict(OFF, pyramid_levels=0, speckle_max_region_size=60), 48, False),
    ("speckles-diff-0.5", 256, 192, 48, dict(OFF, pyramid_levels=1, speckle_max_region_size=200, speckle_max_diff=0.5), 32, False),
    # sizes that leave partial 64x64 speckle tiles on the right and bottom (546 and 407 pixels removed)
    ("speckles-250x170", 250, 170, 48, dict(OFF, pyramid_levels=0, speckle_max_region_size=60), 48, False),
    ("speckles-201x77", 201, 77, 48, dict(OFF, pyramid_levels=0, speckle_max_region_size=60), 48, False),
    ("defaults", 640, 512, 96, dict(), 32, True),
    ("defaults-no-full-pair", 640, 512, 96, dict(), 32, False),
]


def test_refine_blocks_beyond_the_lds_copy():
    """A 480-disparity scene with many depth edges: some 16x16 refine blocks reach more right-image
    columns than the kernel's LDS copy holds (26 blocks by the ground truth's spread) and sample global
    memory instead, in the same launch as blocks that use the copy."""
    ref, got, t = run(640, 480, 480, 91, dict(OFF, pyramid_levels=1), 240, cells=(12, 24))
    assert t["refine_ms"] > 0
    assert (ref["disparity"] > -1e8).mean() > 0.15  # about 64,000 refined pixels compared


@pytest.mark.parametrize("name,w,h,nd,stereo,nd_top,full", STAGES, ids=[c[0] for c in STAGES])
def test_stage_parity(name, w, h, nd, stereo, nd_top, full):
    ref, got, t = run(w, h, nd, 100 + len(name), stereo, nd_top, full)
    if stereo.get("pyramid_levels", 2) > 0:
        assert t["refine_ms"] > 0
    assert (ref["disparity"] > -1e8).mean() > 0.3  # a meaningful number of pixels is compared


def test_subpixel_off():
    pair = speckle_pair(256, 192, 48, 7, min_cells=2, max_cells=3)
    kw = dict(sgm_params(pair, 1, 32), subpixel=False)
    st = dict(OFF, pyramid_levels=1, detail_enabled=True)
    fl = np.repeat(np.repeat(pair.left, 2, 0), 2, 1)
    fr = np.repeat(np.repeat(pair.right, 2, 0), 2, 1)
    ref = O.compute_disparity(pair.left, pair.right, O.SgmParams(**kw), O.StereoParams(**st), fl, fr)
    with sgm.HipStereo.create_stereo(sgm.SgmParams(**kw), sgm.StereoParams(**st), 256, 192) as ctx:
        assert_maps(ref, ctx.compute_disparity(pair.left, pair.right, fl, fr), "subpixel off")


def test_full_size_1280x1024_defaults_with_full_pair():
    """The served frame size with every stage on: SGM at 320x256, two refinement levels, detail
    refinement on the 2560x2048 pair."""
    pair = speckle_pair(2560, 2048, 256, 61, min_cells=3, max_cells=6)
    left, right = O.downsample2(pair.left), O.downsample2(pair.right)
    kw = dict(min_disparity=pair.min_disparity // 8 - 1, num_disparities=32, **SGM)
    ref = O.compute_disparity(left, right, O.SgmParams(**kw), O.StereoParams(), pair.left, pair.right)
    with sgm.HipStereo.create_stereo(sgm.SgmParams(**kw), sgm.StereoParams(), 1280, 1024) as ctx:
        got = ctx.compute_disparity(left, right, pair.left, pair.right)
        again = ctx.compute_disparity(left, right, pair.left, pair.right)
        t = ctx.last_timings()
    assert_maps(ref, got, "1280x1024 defaults")
    assert_maps(got, again, "repeat")
    print("timings", t, flush=True)


def test_raw_frames_with_detail_refinement():
    """compute_disparity_raw: remap at full resolution, downsample2, then the chain with the full
    pair, as MetalStereo::compute_raw runs it (metal_stereo.cpp:448-479)."""
    mw, mh, rw, rh = 640, 512, 660, 530
    maps = rectification_maps(mw, mh, rw, rh, 5)
    raw = speckle_pair(rw, rh, 64, 9)
    kw = dict(min_disparity=raw.min_disparity // 8 - 1, num_disparities=24, **SGM)
    sp = O.StereoParams(speckle_max_region_size=50)
    fl, fr = O.remap(raw.left, maps.left_x, maps.left_y), O.remap(raw.right, maps.right_x, maps.right_y)
    hl, hr = O.downsample2(fl), O.downsample2(fr)
    ref = O.compute_disparity(hl, hr, O.SgmParams(**kw), sp, fl, fr)
    with sgm.HipStereo.create_stereo(sgm.SgmParams(**kw), sgm.StereoParams(speckle_max_region_size=50), mw // 2,
                                     mh // 2) as ctx:
        ctx.set_rectification(maps.left_x, maps.left_y, maps.right_x, maps.right_y, rw, rh)
        got = ctx.compute_disparity_raw(raw.left, raw.right, rectified=True)
        t = ctx.last_timings()
    np.testing.assert_array_equal(got["rect_left"], hl)
    np.testing.assert_array_equal(got["rect_right"], hr)
    assert_maps(ref, got, "raw")
    assert t["rectify_ms"] > 0 and t["refine_ms"] > 0 and t["filter_ms"] > 0


def test_device_entry_points_match_host():
    w, h = 320, 256
    pair = speckle_pair(w, h, 64, 77, min_cells=2, max_cells=4)
    kw = sgm_params(pair, 2, 24)
    fl = np.repeat(np.repeat(pair.left, 2, 0), 2, 1)
    fr = np.repeat(np.repeat(pair.right, 2, 0), 2, 1)
    with sgm.HipStereo.create_stereo(sgm.SgmParams(**kw), sgm.StereoParams(), w, h) as ctx:
        host = ctx.compute_disparity(pair.left, pair.right, fl, fr)
        img, full = sgm.DeviceBuffer(2 * w * h), sgm.DeviceBuffer(8 * w * h)
        img.upload(pair.left, 0)
        img.upload(pair.right, w * h)
        full.upload(fl, 0)
        full.upload(fr, 4 * w * h)
        dd, dc = sgm.DeviceBuffer(4 * w * h), sgm.DeviceBuffer(4 * w * h)
        ctx.compute_disparity_device(img.ptr, img.ptr + w * h, w, w, h, dd.ptr, dc.ptr, full.ptr, full.ptr + 4 * w * h, 2 * w)
        ctx.synchronize()
        dev = {"disparity": dd.download((h, w), np.float32), "confidence": dc.download((h, w), np.float32)}
    assert_maps(host, dev, "device")


def test_plain_context_compute_disparity_is_sgm():
    w, h = 200, 150
    pair = speckle_pair(w, h, 48, 5)
    kw = dict(min_disparity=pair.min_disparity, num_disparities=48, **SGM)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**kw))
    with sgm.HipStereo.create(sgm.SgmParams(**kw), w, h) as ctx:
        got = ctx.compute_disparity(pair.left, pair.right)
    np.testing.assert_array_equal(got["disparity"], ref["disparity"])
    np.testing.assert_array_equal(got["confidence"], (ref["disparity"] > -1e8).astype(np.float32))


def chain_sweep(n, seed):
    """Seeded random compute_disparity configurations: 0..4 pyramid levels, refine search 1..4 and
    ZNCC radius 1..7 (the fixed-radius kernels and the loop form), detail refinement on or off with
    radius 1..16 and steps of 0.25..1 px, each filter on or off with its threshold, speckle sizes
    0..300, with and without the full-resolution pair."""
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        levels = int(rng.integers(0, 5))
        step = float(rng.choice([0.25, 0.5, 1.0]))
        stereo = dict(pyramid_levels=levels, refine_search_radius=int(rng.integers(1, 5)),
                      refine_zncc_radius=int(rng.integers(1, 8)), refine_min_zncc=float(np.float32(rng.uniform(-0.2, 0.9))),
                      refine_reject_edge_winners=bool(rng.random() < 0.5), detail_enabled=bool(rng.random() < 0.6),
                      detail_radius=int(rng.integers(1, 17)), detail_step=step,
                      detail_search=float(step * rng.integers(1, 9)), detail_min_zncc=float(np.float32(rng.uniform(-0.2, 0.9))),
                      filter_occlusion=bool(rng.random() < 0.5),
                      filter_occlusion_tolerance=float(np.float32(rng.uniform(0.25, 3))),
                      filter_spikes=bool(rng.random() < 0.5), filter_spike_threshold=float(np.float32(rng.uniform(0.25, 3))),
                      speckle_max_region_size=int(rng.choice([0, 1, 20, 100, 300])),
                      speckle_max_diff=float(np.float32(rng.uniform(0.25, 2))))
        w, h = int(rng.integers(32, 257)), int(rng.integers(24, 161))
        nd = int(rng.choice([16, 32, 48, 64, 96]))
        cases.append((w, h, nd, 2000 + i, stereo, bool(rng.random() < 0.7), bool(rng.random() < 0.5)))
    return cases


CHAIN_SWEEP = chain_sweep(24, 77)


@pytest.mark.parametrize("w,h,nd,seed,stereo,full,eight", CHAIN_SWEEP,
                         ids=[f"chain{i}-{c[0]}x{c[1]}-L{c[4]['pyramid_levels']}" for i, c in enumerate(CHAIN_SWEEP)])
def test_random_compute_disparity_sweep(w, h, nd, seed, stereo, full, eight):
    pair = speckle_pair(w, h, nd, seed, min_cells=2, max_cells=5)
    levels = stereo["pyramid_levels"]
    kw = dict(min_disparity=pair.min_disparity // (1 << levels) - 1, num_disparities=max(nd >> levels, 1) + 4,
              census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=eight)
    fl = fr = None
    if full:
        fl = np.repeat(np.repeat(pair.left, 2, 0), 2, 1)
        fr = np.repeat(np.repeat(pair.right, 2, 0), 2, 1)
    print(f"chain sweep {w}x{h} D={nd} {kw} {stereo} full={full}", flush=True)
    ref = O.compute_disparity(pair.left, pair.right, O.SgmParams(**kw), O.StereoParams(**stereo), full_left=fl,
                              full_right=fr)
    with sgm.HipStereo.create_stereo(sgm.SgmParams(**kw), sgm.StereoParams(**stereo), w, h) as ctx:
        got = ctx.compute_disparity(pair.left, pair.right, fl, fr)
    assert_maps(ref, got, f"chain sweep {w}x{h}")


def test_wide_frame_compute_disparity():
    """compute_disparity on a frame wider than the row kernel's LDS-staged rows: 9000 columns with one
    pyramid level (SGM at 4500 columns, wide rows), refinement, detail refinement, the occlusion filter
    (one LDS claim per column of the 9000-wide row), spikes and speckles, against the oracle."""
    run(9000, 40, 96, 31, dict(pyramid_levels=1), 64, full=True)
