"""GPU parity: the gfx950 library against the CPU oracle on the same seeded speckle pairs.

Bar: integer outputs (census words, costs, sums, integer disparities, status masks, right
disparities, best/second costs) bit-exact; float outputs (subpixel disparity, median) compared
bit for bit as well, which is stricter than the 1e-4 px tolerance the path allows for them.
"""
import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import speckle_pair

pytestmark = pytest.mark.gpu

FIELDS = ["disparity", "disparity_median", "disparity_int", "status", "right_disparity", "best_cost", "second_cost"]
VOLUMES = ["sum_volume", "cost_volume", "census_left", "census_right"]
SUBPIXEL_TOL = 1e-4  # px, the tolerance stated for the float outputs


def assert_same(ref, got, fields, tag=""):
    for f in fields:
        a, b = np.asarray(ref[f]), np.asarray(got[f])
        assert a.shape == b.shape, (tag, f, a.shape, b.shape)
        if a.dtype == np.float32:
            # Bit-exact is expected; the assertion message reports the worst difference.
            if not np.array_equal(a.view(np.uint32), b.view(np.uint32)):
                va, vb = a > -1e8, b > -1e8
                both = va & vb
                worst = float(np.abs(a[both] - b[both]).max()) if both.any() else 0.0
                assert (va == vb).all() and worst <= SUBPIXEL_TOL, (tag, f, int((va != vb).sum()), worst)
        else:
            if not np.array_equal(a, b):
                idx = np.argwhere(a != b)
                raise AssertionError(f"{tag} {f}: {len(idx)} mismatches, first at {idx[0].tolist()}: "
                                     f"ref {a[tuple(idx[0])]} gpu {b[tuple(idx[0])]}")


def run_case(w, h, nd, seed, volumes=False, **kw):
    pair = speckle_pair(w, h, max(nd, 6), seed) if nd >= 6 else speckle_pair(w, h, 6, seed)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd)
    params.update(kw)
    print(f"case {w}x{h} D={nd} seed={seed} {params}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=volumes)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        got = st.compute(pair.left, pair.right, outputs=FIELDS + (VOLUMES if volumes else []))
    assert_same(ref, got, FIELDS + (VOLUMES if volumes else []), tag=f"{w}x{h}x{nd}")
    return ref, got


SMALL = [
    # (w, h, nd, seed, params)
    (64, 48, 16, 1, dict()),
    (64, 48, 16, 2, dict(eight_paths=True)),
    (64, 48, 24, 3, dict(eight_paths=True, census_radius_x=4, census_radius_y=3, p1=10, p2=120)),
    (96, 64, 32, 4, dict(adaptive_p2=False, slant_steps=False)),
    (96, 64, 32, 5, dict(adaptive_p2=False, p1=12, p2=7, eight_paths=True)),
    (80, 40, 7, 6, dict(census_radius_x=1, census_radius_y=1)),
    (70, 50, 64, 7, dict(eight_paths=True, census_radius_x=3, census_radius_y=1, p1=5, p2=60)),
    (40, 30, 16, 8, dict(min_disparity=-6)),
    (40, 30, 16, 9, dict(min_disparity=-20, eight_paths=True)),
    (20, 12, 64, 10, dict(eight_paths=True)),
    (33, 17, 40, 11, dict(eight_paths=True, uniqueness=0.7, lr_max_diff=0)),
    (50, 45, 20, 12, dict(subpixel=False, eight_paths=True)),
    (48, 36, 16, 13, dict(census_radius_x=0, census_radius_y=0)),
    (64, 48, 16, 14, dict(p1=0, p2=0, adaptive_p2=False, eight_paths=True)),
]


@pytest.mark.parametrize("w,h,nd,seed,kw", SMALL, ids=[f"{c[0]}x{c[1]}x{c[2]}-s{c[3]}" for c in SMALL])
def test_small_pairs_bit_exact_with_volumes(w, h, nd, seed, kw):
    run_case(w, h, nd, seed, volumes=True, **kw)


EDGES = [(1, 1, 8), (1, 40, 16), (40, 1, 16), (2, 2, 4), (5, 3, 32)]


@pytest.mark.parametrize("w,h,nd", EDGES, ids=[f"{w}x{h}x{nd}" for w, h, nd in EDGES])
def test_degenerate_sizes(w, h, nd):
    rng = np.random.default_rng(w * 100 + h)
    left = rng.integers(0, 256, (h, w), dtype=np.uint8)
    right = rng.integers(0, 256, (h, w), dtype=np.uint8)
    params = dict(num_disparities=nd, eight_paths=True, census_radius_x=1, census_radius_y=1)
    print(f"degenerate {w}x{h} D={nd}", flush=True)
    ref = O.sgm(left, right, O.SgmParams(**params), volumes=True)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        got = st.compute(left, right, outputs=FIELDS + VOLUMES)
    assert_same(ref, got, FIELDS + VOLUMES, tag=f"{w}x{h}")


LANES = [96, 128, 192, 256, 384, 512]


@pytest.mark.parametrize("nd", LANES)
def test_every_lane_layout(nd):
    """One case per disparities-per-lane layout above 64 (8 with 32 padded disparities, 8, 12, 16, 24,
    32)."""
    run_case(nd + 40, 24, nd, 20 + nd, volumes=True, eight_paths=True, census_radius_x=4, census_radius_y=3,
             p1=10, p2=120)


PADDED_LARGE = [(640, 512, 24, False), (800, 600, 96, True), (520, 510, 32, True)]


@pytest.mark.parametrize("w,h,nd,eight", PADDED_LARGE, ids=[f"{c[0]}x{c[1]}x{c[2]}" for c in PADDED_LARGE])
def test_padded_layouts_on_large_pairs(w, h, nd, eight):
    """Disparity counts whose narrow lane layouts (2 and 6 per lane) are not used on pairs above 256K
    pixels: D <= 32 runs 4 per lane and D = 96 runs 8 per lane, with the padded disparities masked;
    with volumes, in both cost modes."""
    ref, _ = run_case(w, h, nd, 60 + nd, volumes=True, eight_paths=eight, census_radius_x=4, census_radius_y=3,
                      p1=10, p2=120)
    pair = speckle_pair(w, h, nd, 60 + nd)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd, eight_paths=eight, census_radius_x=4,
                  census_radius_y=3, p1=10, p2=120)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        st.set_option(sgm.OPT_COST_VOLUME, 0)
        got = st.compute(pair.left, pair.right, outputs=FIELDS)
    assert_same(ref, got, FIELDS, tag=f"{w}x{h}x{nd} recomputed costs")


PER_PAIR_GROUP = [
    # (w, h, nd, seed, cost_volume, recompute_dirs, params): lanes per scanline left to the library
    (320, 256, 247, 81, 1, 0, dict()),                      # 4 paths, 256 padded: 32 lanes, 8 lines
    (300, 200, 256, 82, 1, 0, dict(eight_paths=True)),     # 8 paths, 256 disparities, small: 32 lanes
    (256, 200, 384, 83, 1, 0, dict()),                      # 4 paths, 384: 32 lanes, 4 lines
    (400, 300, 128, 84, 0, 0, dict(eight_paths=True)),     # recomputed costs: 32 lanes
    (400, 300, 128, 85, 1, 4, dict(eight_paths=True)),     # half the directions recompute: 32 lanes
    (400, 300, 192, 86, 0, 0, dict(eight_paths=True)),     # 12 per lane: 16 lanes
    (600, 40, 512, 87, 1, 0, dict()),                       # 32 per lane: 64 lanes
    (300, 200, 256, 88, 0, 0, dict()),                      # 16 per lane, recomputed costs: 64 lanes
]


@pytest.mark.parametrize("w,h,nd,seed,cv,rd,kw", PER_PAIR_GROUP,
                         ids=[f"{c[0]}x{c[1]}x{c[2]}-cv{c[4]}-rd{c[5]}" for c in PER_PAIR_GROUP])
def test_lanes_per_scanline_chosen_per_pair(w, h, nd, seed, cv, rd, kw):
    """With the scanline width and lines per block left open, each branch of the per-pair choice
    (64 lanes for 32 disparities per lane and for recomputed costs at 16; 32 lanes for small 4-path
    pairs, for recomputed costs and for wide ranges on small 8-path pairs; 16 where the layout holds 12
    disparities per lane) gives the oracle's bytes, and so does the 16-lane width forced on the same
    context."""
    pair = speckle_pair(w, h, min(nd, 256), seed)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd, census_radius_x=4, census_radius_y=3,
                  p1=10, p2=120)
    params.update(kw)
    volumes = w * h * nd <= 20_000_000
    print(f"per-pair lanes {w}x{h} D={nd} cv={cv} rd={rd} {kw}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=volumes)
    fields = FIELDS + (VOLUMES if volumes else [])
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        st.set_option(sgm.OPT_COST_VOLUME, cv)
        st.set_option(sgm.OPT_RECOMPUTE_DIRS, rd)
        got = st.compute(pair.left, pair.right, outputs=fields)
        assert_same(ref, got, fields, tag=f"{w}x{h}x{nd} lanes chosen per pair")
        st.set_option(sgm.OPT_PATH_GROUP, 16)
        got = st.compute(pair.left, pair.right, outputs=FIELDS)
        assert_same(ref, got, FIELDS, tag=f"{w}x{h}x{nd} 16 lanes")
        st.set_option(sgm.OPT_PATH_GROUP, 0)
        got = st.compute(pair.left, pair.right, outputs=FIELDS)
        assert_same(ref, got, FIELDS, tag=f"{w}x{h}x{nd} lanes chosen per pair again")


ROW_BLOCKS = [
    # (w, h, nd, seed, params): large row blocks of 1024 threads, or 512 where 1024 would spill
    (320, 256, 247, 91, dict()),                                       # 16 per lane, 4 paths: 1024
    (300, 120, 256, 92, dict(eight_paths=True)),                       # 16 per lane, 8 paths: 512
    (200, 64, 512, 93, dict()),                                        # 32 per lane: 512
    (257, 200, 96, 94, dict(eight_paths=True, min_disparity=-9)),      # 8 per lane, padded
    (40, 300, 32, 95, dict(eight_paths=True)),                         # fewer columns than groups, 300 rows
    (4096, 20, 64, 96, dict(eight_paths=True)),                        # the widest staged rows
    (4200, 8, 64, 97, dict(eight_paths=True)),                         # wider rows run 256 threads whatever is asked
]


@pytest.mark.parametrize("w,h,nd,seed,kw", ROW_BLOCKS, ids=[f"{c[0]}x{c[1]}x{c[2]}" for c in ROW_BLOCKS])
def test_row_block_sizes(w, h, nd, seed, kw):
    """Every threads-per-row-block setting (chosen per pair, 256, the large block) gives the oracle's
    bytes, sums included, on one context."""
    pair = speckle_pair(w, h, min(nd, 256), seed)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd, census_radius_x=4, census_radius_y=3,
                  p1=10, p2=120)
    params.update(kw)
    volumes = w * h * nd <= 20_000_000
    fields = FIELDS + (VOLUMES if volumes else [])
    print(f"row blocks {w}x{h} D={nd} {kw}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=volumes)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        for rt in (0, 256, 1024, 0):
            st.set_option(sgm.OPT_ROW_THREADS, rt)
            got = st.compute(pair.left, pair.right, outputs=fields)
            assert_same(ref, got, fields, tag=f"{w}x{h}x{nd} row threads {rt}")


def test_layout_changes_with_the_pair_size_on_one_context():
    """D <= 32 runs 2 disparities per lane on pairs of at most 256K pixels and 4 per lane above, chosen
    per call: one context sized for the larger layout computes pairs on both sides of the boundary
    (512x512 and 513x512), a small one and its maximum size, each against the oracle."""
    params = dict(min_disparity=4, num_disparities=24, eight_paths=True, census_radius_x=4, census_radius_y=3,
                  p1=10, p2=120)
    with sgm.HipStereo.create(sgm.SgmParams(**params), 600, 600) as st:
        for w, h, seed in ((512, 512, 71), (513, 512, 72), (200, 150, 73), (600, 600, 74), (512, 512, 75)):
            pair = speckle_pair(w, h, 24, seed)
            print(f"layout boundary {w}x{h}", flush=True)
            ref = O.sgm(pair.left, pair.right, O.SgmParams(**params))
            got = st.compute(pair.left, pair.right, outputs=FIELDS)
            assert_same(ref, got, FIELDS, tag=f"{w}x{h}")
            assert (ref["disparity"] > -1e8).mean() > 0.3


def test_wide_lanes_second_best_full_window():
    """512 disparities (32 per lane) on an image whose right part sees every disparity, so the row
    kernel's full-window branch picks best disparities in every lane slot (1,051 pixels in slots 30
    and 31, where the second-best exclusion reaches past bit 31 of a 32-bit mask), with a strict
    uniqueness ratio so that the second-best cost decides many pixels."""
    run_case(1100, 48, 512, 77, census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=True,
             uniqueness=0.95)


def test_staged_row_width_4096_and_first_wide_row_4097():
    """The widest row whose left results the row kernel stages in LDS (4096 columns, 15 bytes each) and
    one column more, where the kernel keeps only the right WTA keys in LDS and writes the left results
    straight to the outputs; the largest disparity count, 8 paths and the 9x7 census, with volumes,
    then in both cost modes."""
    for w in (4096, 4097):
        h, nd = 24, 512
        pair = speckle_pair(w, h, nd, 93)
        params = dict(min_disparity=pair.min_disparity, num_disparities=nd, eight_paths=True, census_radius_x=4,
                      census_radius_y=3, p1=10, p2=120)
        print(f"width {w}x{h} D={nd} {params}", flush=True)
        ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=True)
        with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
            got = st.compute(pair.left, pair.right, outputs=FIELDS + VOLUMES)
            assert_same(ref, got, FIELDS + VOLUMES, tag=f"{w}x{h}x{nd}")
            for cv in (1, 0):
                st.set_option(sgm.OPT_COST_VOLUME, cv)
                got = st.compute(pair.left, pair.right, outputs=FIELDS)
                assert_same(ref, got, FIELDS, tag=f"{w}x{h}x{nd} cost volume {cv}")
        assert (ref["disparity"] > -1e8).mean() > 0.5


def test_maximum_width_32768():
    """The widest row the library accepts (32768 columns: 4 bytes of LDS per column in the row and
    occlusion kernels), both cost modes and both path-lane widths, a narrower image on the same
    context, and one column more refused at create time."""
    w, h, nd = 32768, 8, 64
    pair = speckle_pair(w, h, nd, 94)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd, eight_paths=True, census_radius_x=4,
                  census_radius_y=3, p1=10, p2=120)
    print(f"maximum width {w}x{h} D={nd}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params))
    small = speckle_pair(5000, 6, nd, 95)
    ref_small = O.sgm(small.left, small.right, O.SgmParams(**dict(params, min_disparity=small.min_disparity)))
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        for cv in (1, 0):
            for group in (16, 32):
                st.set_option(sgm.OPT_COST_VOLUME, cv)
                st.set_option(sgm.OPT_PATH_GROUP, group)
                got = st.compute(pair.left, pair.right, outputs=FIELDS)
                assert_same(ref, got, FIELDS, tag=f"{w}x{h}x{nd} cost volume {cv} lanes {group}")
    with sgm.HipStereo.create(sgm.SgmParams(**dict(params, min_disparity=small.min_disparity)), w, h) as st:
        got = st.compute(small.left, small.right, outputs=FIELDS)
        assert_same(ref_small, got, FIELDS, tag="5000x6 on a 32768-wide context")
    with pytest.raises(sgm.SgmError) as e:
        sgm.HipStereo.create(sgm.SgmParams(**params), w + 1, h)
    assert "max_width" in str(e.value)


def test_whole_640x400x64_four_paths_census_9x7():
    run_case(640, 400, 64, 1, census_radius_x=4, census_radius_y=3, p1=10, p2=120)


def test_whole_1280x1024x128_eight_paths_census_9x7():
    run_case(1280, 1024, 128, 2, census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=True)


def test_repeated_frames_and_context_reuse_leak_no_state():
    """test_metal_stereo.cpp:152-158: repeated frames give identical disparities; here also one
    context reused for pairs of different sizes up to its maximum (the generator's disparity range
    depends only on D, so min_disparity is shared)."""
    base = dict(num_disparities=32, eight_paths=True, census_radius_x=4, census_radius_y=3, p1=10, p2=120)
    min_d = speckle_pair(16, 16, 32, 0).min_disparity
    params = dict(base, min_disparity=min_d)
    with sgm.HipStereo.create(sgm.SgmParams(**params), 160, 120) as st:
        for (w, h, seed) in [(160, 120, 30), (64, 48, 31), (160, 120, 30), (97, 33, 32)]:
            pair = speckle_pair(w, h, 32, seed)
            assert pair.min_disparity == min_d
            a = st.compute(pair.left, pair.right, outputs=FIELDS)
            b = st.compute(pair.left, pair.right, outputs=FIELDS)
            ref = O.sgm(pair.left, pair.right, O.SgmParams(**params))
            assert_same(ref, a, FIELDS, tag=f"reuse {w}x{h}")
            assert_same(a, b, FIELDS, tag=f"repeat {w}x{h}")


def test_device_entry_point_matches_host_entry_point():
    w, h, nd = 200, 150, 48
    pair = speckle_pair(w, h, nd, 40)
    params = sgm.SgmParams(min_disparity=pair.min_disparity, num_disparities=nd, eight_paths=True)
    with sgm.HipStereo.create(params, w, h) as st:
        host = st.compute(pair.left, pair.right, outputs=FIELDS)
        img = sgm.DeviceBuffer(2 * w * h)
        img.upload(pair.left, 0)
        img.upload(pair.right, w * h)
        outs = {f: sgm.DeviceBuffer(w * h * np.dtype(sgm.OUTPUT_DTYPES[f]).itemsize) for f in FIELDS}
        st.compute_device(img.ptr, img.ptr + w * h, w, w, h, {f: b.ptr for f, b in outs.items()})
        st.synchronize()
        dev = {f: b.download((h, w), sgm.OUTPUT_DTYPES[f]) for f, b in outs.items()}
        t = st.last_timings()
    assert_same(host, dev, FIELDS, tag="device")
    assert t["gpu_ms"] > 0 and t["paths_ms"] > 0


def test_size_errors_on_device():
    with sgm.HipStereo.create(sgm.SgmParams(num_disparities=16), 32, 32) as st:
        with pytest.raises(sgm.SgmError) as e:
            st.compute(np.zeros((40, 32), np.uint8), np.zeros((40, 32), np.uint8))
        assert e.value.errc == "invalid_argument"


CV_CASES = [(64, 48, 16, 51, dict()), (64, 48, 24, 52, dict(eight_paths=True, census_radius_x=4, census_radius_y=3,
            p1=10, p2=120)), (40, 30, 16, 53, dict(min_disparity=-20, eight_paths=True)), (20, 12, 64, 54, dict()),
            (230, 40, 192, 55, dict(eight_paths=True)), (136, 24, 96, 56, dict(eight_paths=True)),
            (552, 24, 512, 57, dict(eight_paths=True)), (1, 40, 16, 58, dict(eight_paths=True))]


@pytest.mark.parametrize("w,h,nd,seed,kw", CV_CASES, ids=[f"cv-{c[0]}x{c[1]}x{c[2]}" for c in CV_CASES])
def test_cost_volume_mode_bit_exact(w, h, nd, seed, kw):
    """The cost-volume mode (costs written once as bytes) gives the same bytes as recomputation."""
    pair = speckle_pair(w, h, max(nd, 6), seed)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd)
    params.update(kw)
    print(f"cost-volume case {w}x{h} D={nd} {params}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params))
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        st.set_option(sgm.OPT_COST_VOLUME, 1)
        st.set_option(sgm.OPT_PATH_LINES, 4)
        got = st.compute(pair.left, pair.right, outputs=FIELDS)
        st.set_option(sgm.OPT_COST_VOLUME, 0)
        st.set_option(sgm.OPT_PATH_LINES, 8)
        again = st.compute(pair.left, pair.right, outputs=FIELDS)
    assert_same(ref, got, FIELDS, tag=f"cv {w}x{h}x{nd}")
    assert_same(ref, again, FIELDS, tag=f"recompute after cv {w}x{h}x{nd}")


G32_CASES = [(96, 64, 64, 81, dict()), (80, 48, 100, 82, dict(eight_paths=True)), (64, 40, 128, 83, dict(eight_paths=True)),
             (120, 30, 192, 84, dict(eight_paths=True, min_disparity=-10)), (300, 20, 256, 85, dict(eight_paths=True)),
             (420, 16, 384, 86, dict(eight_paths=True)), (560, 12, 512, 87, dict(eight_paths=True)),
             (40, 30, 56, 88, dict(min_disparity=-20, eight_paths=True, census_radius_x=4, census_radius_y=3, p1=10, p2=120))]


@pytest.mark.parametrize("cv", [0, 1], ids=["recompute", "volume"])
@pytest.mark.parametrize("w,h,nd,seed,kw", G32_CASES, ids=[f"g32-{c[0]}x{c[1]}x{c[2]}" for c in G32_CASES])
def test_32_lane_scanlines_bit_exact(w, h, nd, seed, kw, cv):
    """Scanlines spread over 32 lanes (two DPP rows, exchanging across the row boundary) give the
    same bytes as the oracle, for every lane layout that allows them, padded or not."""
    pair = speckle_pair(w, h, max(nd, 6), seed)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd)
    params.update(kw)
    print(f"g32 {w}x{h} D={nd} cv={cv} {params}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=True)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        st.set_option(sgm.OPT_COST_VOLUME, cv)
        st.set_option(sgm.OPT_PATH_GROUP, 32)
        for lines in (16, 8, 4):
            st.set_option(sgm.OPT_PATH_LINES, lines)
            got = st.compute(pair.left, pair.right, outputs=FIELDS + VOLUMES)
            assert_same(ref, got, FIELDS + VOLUMES, tag=f"g32 {w}x{h}x{nd} lines {lines}")


G64_CASES = [(300, 20, 256, 101, dict(eight_paths=True)), (280, 24, 247, 102, dict()),
             (560, 12, 512, 103, dict(eight_paths=True)), (520, 10, 450, 104, dict(slant_steps=False)),
             (200, 40, 200, 105, dict(eight_paths=True, min_disparity=-30, census_radius_x=4, census_radius_y=3,
                                      p1=10, p2=120)),
             (96, 64, 128, 106, dict(eight_paths=True)),  # 8 per 16-lane lane: runs 32 lanes
             (120, 30, 192, 107, dict())]                 # 12 per 16-lane lane: runs 32 lanes (6 per lane)


@pytest.mark.parametrize("cv", [0, 1], ids=["recompute", "volume"])
@pytest.mark.parametrize("w,h,nd,seed,kw", G64_CASES, ids=[f"g64-{c[0]}x{c[1]}x{c[2]}" for c in G64_CASES])
def test_64_lane_scanlines_bit_exact(w, h, nd, seed, kw, cv):
    """Scanlines spread over the whole wave (64 lanes, exchanging across the DPP rows and the wave's
    halves) give the same bytes as the oracle for the layouts that allow them (16 and 32 disparities per
    16-lane lane, padded or not), whatever lines per block are asked for; other layouts run 32 or 16
    lanes instead."""
    pair = speckle_pair(w, h, min(nd, 256), seed)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd)
    params.update(kw)
    print(f"g64 {w}x{h} D={nd} cv={cv} {params}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=True)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        st.set_option(sgm.OPT_COST_VOLUME, cv)
        st.set_option(sgm.OPT_PATH_GROUP, 64)
        for lines in (0, 4, 16):
            st.set_option(sgm.OPT_PATH_LINES, lines)
            got = st.compute(pair.left, pair.right, outputs=FIELDS + VOLUMES)
            assert_same(ref, got, FIELDS + VOLUMES, tag=f"g64 {w}x{h}x{nd} lines {lines}")


def sweep_cases(n, seed):
    """Seeded random cases over the whole accepted parameter space and every tuning knob: sizes
    1..300 x 1..60, 1..512 disparities, census windows 1x1..9x7, P1/P2 up to the byte bound (P1 > P2
    included), min_disparity -80..80, uniqueness 0.5..1, L-R tolerance -1..3, each flag on or off."""
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        rx, ry = int(rng.integers(0, 5)), int(rng.integers(0, 4))
        cells = (2 * rx + 1) * (2 * ry + 1)
        nd = int(rng.choice([1, 2, 3, 5, 16, 17, 31, 64, 65, 100, 128, 130, 192, 200, 256, 300, 384, 511, 512]))
        params = dict(min_disparity=int(rng.integers(-80, 81)), num_disparities=nd, census_radius_x=rx,
                      census_radius_y=ry, p1=int(rng.integers(0, 255 - cells)), p2=int(rng.integers(0, 256 - cells)),
                      adaptive_p2=bool(rng.random() < 0.5), eight_paths=bool(rng.random() < 0.6),
                      slant_steps=bool(rng.random() < 0.5), uniqueness=float(np.float32(rng.uniform(0.5, 1.0))),
                      lr_max_diff=int(rng.integers(-1, 4)), subpixel=bool(rng.random() < 0.8))
        knobs = {sgm.OPT_COST_VOLUME: int(rng.integers(0, 2)), sgm.OPT_PATH_LINES: int(rng.choice([0, 4, 8, 16])),
                 sgm.OPT_PATH_GROUP: int(rng.choice([0, 16, 32, 64])), sgm.OPT_RECOMPUTE_DIRS: int(rng.integers(0, 9)),
                 sgm.OPT_ROW_THREADS: int(rng.choice([0, 256, 1024]))}
        cases.append((int(rng.integers(1, 301)), int(rng.integers(1, 61)), 1000 + i, params, knobs))
    return cases


SWEEP = sweep_cases(48, 2026)


@pytest.mark.parametrize("w,h,seed,params,knobs", SWEEP, ids=[f"sweep{i}-{c[0]}x{c[1]}x{c[3]['num_disparities']}"
                                                               for i, c in enumerate(SWEEP)])
def test_random_parameter_sweep(w, h, seed, params, knobs):
    pair = speckle_pair(w, h, max(params["num_disparities"], 6), seed)
    print(f"sweep {w}x{h} {params} knobs {knobs}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=True)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        for k, v in knobs.items():
            st.set_option(k, v)
        got = st.compute(pair.left, pair.right, outputs=FIELDS + VOLUMES)
    assert_same(ref, got, FIELDS + VOLUMES, tag=f"sweep {w}x{h}")


VIEW_CASES = [(1, 16), (1, 32), (0, 16), (0, 32)]  # (cost volume, lanes per scanline)


@pytest.mark.parametrize("cv,group", VIEW_CASES, ids=[f"cv{cv}-g{g}" for cv, g in VIEW_CASES])
def test_view_bases_with_bit31_set(cv, group):
    """Regression test of the buffer views' base address: the census words, step penalties and cost
    volume are placed where the low 32-bit word of the address has bit 31 set (SGM_OPT_VIEW_BASE_HIGH),
    so that a sign extension of that word while building a view would read the wrong memory."""
    w, h, nd = 200, 96, 128
    pair = speckle_pair(w, h, nd, 4242)
    params = dict(min_disparity=pair.min_disparity, num_disparities=nd, eight_paths=True, census_radius_x=4,
                  census_radius_y=3, p1=10, p2=120)
    print(f"view bases bit 31: cv={cv} lanes={group}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=True)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        st.set_option(sgm.OPT_VIEW_BASE_HIGH, 1)
        st.set_option(sgm.OPT_COST_VOLUME, cv)
        st.set_option(sgm.OPT_PATH_GROUP, group)
        for _ in range(2):
            got = st.compute(pair.left, pair.right, outputs=FIELDS + VOLUMES)
            assert_same(ref, got, FIELDS + VOLUMES, tag=f"high view bases cv={cv} lanes={group}")


WIDE_COSTS = [
    # (w, h, nd, seed, params): each needs u16 path costs (cells + max step penalty > 255, or P1 > 32767)
    (300, 60, 64, 41, dict(eight_paths=True, census_radius_x=4, census_radius_y=3, p1=200, p2=100)),
    (96, 64, 32, 42, dict(adaptive_p2=False, p1=40000, p2=30)),
    (3000, 12, 64, 43, dict(eight_paths=True, census_radius_x=4, census_radius_y=3, p1=100000, p2=120)),
    (200, 100, 100, 44, dict(eight_paths=True, p1=500, p2=60, min_disparity=-7)),
    (150, 40, 48, 45, dict(census_radius_x=2, census_radius_y=2, p1=1 << 30, p2=230, slant_steps=False)),
    (5000, 10, 64, 46, dict(eight_paths=True, census_radius_x=4, census_radius_y=3, p1=70000, p2=120)),
]


@pytest.mark.parametrize("w,h,nd,seed,kw", WIDE_COSTS, ids=[f"{c[0]}x{c[1]}x{c[2]}-p1_{c[4]['p1']}" for c in WIDE_COSTS])
def test_u16_path_costs_bit_exact(w, h, nd, seed, kw):
    """Penalties beyond what byte path costs hold (MetalStereo accepts any P1 with cells + P2 <= 255):
    u16 path costs and step penalties, saturating candidate sums, 16-bit wrap of path costs and sums as
    the reference's u16 casts, with volumes, in both cost modes; the long rows with P1 >= 70000 carry
    path costs past 2^16."""
    ref, _ = run_case(w, h, nd, seed, volumes=w * h * nd <= 2_000_000, **kw)
    pair = speckle_pair(w, h, max(nd, 6), seed)
    params = dict(min_disparity=kw.get("min_disparity", pair.min_disparity), num_disparities=nd, **{k: v for k, v in kw.items() if k != "min_disparity"})
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        for cv, group in ((0, 16), (1, 32), (1, 64), (0, 0)):  # u16 path costs run 16 lanes whatever is asked
            st.set_option(sgm.OPT_COST_VOLUME, cv)
            st.set_option(sgm.OPT_PATH_GROUP, group)
            got = st.compute(pair.left, pair.right, outputs=FIELDS)
            assert_same(ref, got, FIELDS, tag=f"u16 path costs {w}x{h}x{nd} cv={cv} lanes={group}")


def test_u16_path_costs_in_a_batch_and_served():
    params = dict(eight_paths=True, census_radius_x=4, census_radius_y=3, p1=300, p2=120)
    sizes = [(200, 80, 64), (150, 60, 96), (64, 48, 128)]
    pairs = [speckle_pair(w, h, d, 7000 + i) for i, (w, h, d) in enumerate(sizes)]
    refs = [O.sgm(p.left, p.right, O.SgmParams(min_disparity=p.min_disparity, num_disparities=d, **params))
            for p, (_, _, d) in zip(pairs, sizes)]
    specs = [sgm.PairSpec(p.left, p.right, p.min_disparity, d, outputs=tuple(FIELDS)) for p, (_, _, d) in zip(pairs, sizes)]
    b = sgm.Batch(sgm.SgmParams(**params), specs, tuning=dict(cost_volume=1, path_lines=8, branches=1, path_group=32))
    try:
        for mode in ("graph", "plain"):
            b.clear_outputs(0xEE)
            b.run() if mode == "graph" else b.run_plain()
            b.synchronize()
            for i in range(len(sizes)):
                assert_same(refs[i], b.results(i), FIELDS, tag=f"u16 batch {mode} pair {i}")
    finally:
        b.close()
    with sgm.Server(sgm.SgmParams(**params), 200, 80, 128, slots=2) as sv:
        for i, (p, (_, _, d)) in enumerate(zip(pairs, sizes)):
            sv.submit(p.left, p.right, p.min_disparity, d, outputs=FIELDS, tag=i)
        got = {}
        import time
        t0 = time.time()
        while len(got) < len(sizes) and time.time() - t0 < 60:
            for c in sv.poll(8, timeout_us=100000):
                got[c["tag"]] = {f: a.copy() for f, a in c["outputs"].items()}
                sv.release(c)
    for i in range(len(sizes)):
        assert_same(refs[i], got[i], FIELDS, tag=f"u16 served request {i}")


def test_u16_path_costs_wrap_on_long_rows():
    """Random-noise rows shifted by a constant disparity: every wrong disparity costs about 31 at each
    pixel, and with P1 = 100000 (adaptive P2: step penalty 100001) its horizontal path cost keeps its
    own previous value plus that cost while the line minimum stays near 0, so it passes 2^16 after
    about 2,100 pixels of a 5000-pixel row and wraps; the sums wrap with it. Bit-exact in both cost
    modes (the guard below confirms, from the oracle's own cost volume, that the unwrapped
    left-to-right path cost exceeds 2^16)."""
    w, h, nd, shift = 5000, 8, 64, 20
    rng = np.random.default_rng(4711)
    right = rng.integers(0, 256, (h, w + shift), dtype=np.uint8)
    left = np.ascontiguousarray(right[:, :w])
    right = np.ascontiguousarray(right[:, shift:shift + w])
    kw = dict(min_disparity=0, num_disparities=nd, eight_paths=True, census_radius_x=4, census_radius_y=3,
              p1=100000, p2=120)
    ref = O.sgm(left, right, O.SgmParams(**kw), volumes=True)
    C = ref["cost_volume"][h // 2].astype(np.int64)  # an interior row (census rows 3..h-4)
    L, mx = C[0].copy(), 0
    for x in range(1, w):  # rows of constant-free noise: p2 = max(P1 + 1, ...) = P1 + 1
        pm = L.min()
        best = np.minimum(L, pm + kw["p1"] + 1)
        L = C[x] + best - pm
        mx = max(mx, int(L.max()))
    assert mx > 65535, mx
    with sgm.HipStereo.create(sgm.SgmParams(**kw), w, h) as st:
        for cv in (1, 0):
            st.set_option(sgm.OPT_COST_VOLUME, cv)
            got = st.compute(left, right, outputs=FIELDS + ["sum_volume"])
            assert_same(ref, got, FIELDS + ["sum_volume"], tag=f"wrapping u16 path costs cv={cv}")
