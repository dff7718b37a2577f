"""Batched SGM: pairs of uneven sizes and disparity ranges in one graph launch give, pair by pair,
the oracle's bytes; the graph replays and the plain per-pair launches agree with each other."""
import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import speckle_pair
from test_gpu_parity import FIELDS, assert_same

pytestmark = pytest.mark.gpu

BASE = dict(census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=True)


def make_specs(sizes, seed0):
    specs, refs = [], []
    for i, (w, h, nd) in enumerate(sizes):
        pair = speckle_pair(w, h, nd, seed0 + i)
        specs.append(sgm.PairSpec(pair.left, pair.right, pair.min_disparity, nd, outputs=tuple(FIELDS)))
        p = O.SgmParams(min_disparity=pair.min_disparity, num_disparities=nd, **BASE)
        refs.append(O.sgm(pair.left, pair.right, p))
    return specs, refs


def test_uneven_batch_bit_exact_and_replayable():
    sizes = [(96, 64, 64), (130, 70, 96), (64, 48, 128), (200, 40, 192), (77, 91, 64), (150, 110, 128), (40, 30, 24)]
    specs, refs = make_specs(sizes, 1000)
    print(f"batch of {len(sizes)}", flush=True)
    b = sgm.Batch(sgm.SgmParams(**BASE), specs)
    try:
        b.run()
        b.synchronize()
        for i in range(len(sizes)):
            assert_same(refs[i], b.results(i), FIELDS, tag=f"batch pair {i} {sizes[i]}")
        # Replays after clearing the outputs, and the plain per-pair launches, give the same bytes.
        for mode in ("graph", "plain", "graph"):
            b.clear_outputs(0xEE)
            b.run() if mode == "graph" else b.run_plain()
            b.synchronize()
            for i in range(len(sizes)):
                assert_same(refs[i], b.results(i), FIELDS, tag=f"{mode} pair {i}")
    finally:
        b.close()


def test_single_pair_batch():
    specs, refs = make_specs([(120, 80, 48)], 2000)
    b = sgm.Batch(sgm.SgmParams(**BASE), specs)
    try:
        b.run()
        b.synchronize()
        assert_same(refs[0], b.results(0), FIELDS, tag="single")
    finally:
        b.close()


@pytest.mark.parametrize("tuning", [dict(cost_volume=0, path_lines=4, branches=0, path_group=16, sweep_band=0),
                                    dict(cost_volume=1, path_lines=8, branches=2, path_group=32, sweep_band=0),
                                    dict(cost_volume=1, path_lines=16, branches=1, path_group=16, sweep_band=0),
                                    dict(cost_volume=0, path_lines=8, branches=1, path_group=32, sweep_band=0),
                                    dict(cost_volume=1, path_lines=8, branches=1, path_group=16, sweep_band=32, sweep_tile=64),
                                    dict(cost_volume=1, path_lines=16, branches=0, path_group=16, sweep_band=5, sweep_tile=16),
                                    dict(cost_volume=1, path_lines=4, branches=2, path_group=32, sweep_band=1, sweep_tile=48),
                                    dict(cost_volume=1, path_lines=8, branches=1, path_group=16, sweep_band=200, sweep_tile=32)],
                         ids=["recompute-4-onestream", "volume-8-split-32lanes", "volume-16-perclass",
                              "recompute-8-perclass-32lanes", "sweeps-32rows-64cols", "sweeps-5rows-16cols-onestream",
                              "sweeps-1row-48cols-split", "sweeps-one-band-32cols"])
def test_every_plan_gives_the_same_bytes(tuning):
    """Line and band-sweep plans: every pair equals the oracle, the graph and the plain launches. The
    sizes hold tiles and bands that do not divide the image, padded lanes (D = 24, 96, 192) and a
    pair taller than one band of 64."""
    sizes = [(96, 64, 64), (130, 70, 96), (64, 48, 128), (77, 91, 64), (40, 30, 24), (33, 130, 192)]
    specs, refs = make_specs(sizes, 3000)
    b = sgm.Batch(sgm.SgmParams(**BASE), specs, tuning=tuning)
    try:
        plan = b.plan()
        assert all(plan[k] == v for k, v in tuning.items()) and plan["candidates_timed"] == 0
        for mode in ("graph", "plain"):
            b.clear_outputs(0xEE)
            b.run() if mode == "graph" else b.run_plain()
            b.synchronize()
            for i in range(len(sizes)):
                assert_same(refs[i], b.results(i), FIELDS, tag=f"{tuning} {mode} pair {i}")
    finally:
        b.close()


@pytest.mark.parametrize("slant,adaptive", [(True, True), (False, False)], ids=["slant-adaptive", "plain-steps"])
def test_band_sweep_sum_volumes(slant, adaptive):
    """The sum volume S under band sweeps (the swept u16 sums of six directions plus the two horizontal
    path costs, added by the row kernel) equals the oracle's for every pixel and disparity, on pairs
    whose tiles and bands do not divide them, with negative and positive disparity offsets."""
    params = dict(BASE, slant_steps=slant, adaptive_p2=adaptive)
    sizes = [(150, 75, 64, -5), (70, 140, 40, 3), (23, 9, 16, 0)]
    specs, refs = [], []
    for i, (w, h, nd, min_d) in enumerate(sizes):
        pair = speckle_pair(w, h, nd, 7100 + i)
        specs.append(sgm.PairSpec(pair.left, pair.right, min_d, nd, outputs=tuple(FIELDS) + ("sum_volume",)))
        refs.append(O.sgm(pair.left, pair.right, O.SgmParams(min_disparity=min_d, num_disparities=nd, **params),
                          volumes=True))
    b = sgm.Batch(sgm.SgmParams(**params), specs,
                  tuning=dict(cost_volume=1, path_lines=8, branches=1, path_group=16, sweep_band=12, sweep_tile=32))
    try:
        for rep in range(2):
            b.clear_outputs(0xEE)
            b.run()
            b.synchronize()
            for i in range(len(sizes)):
                assert_same(refs[i], b.results(i), FIELDS + ["sum_volume"], tag=f"sweep rep {rep} pair {i} {sizes[i]}")
    finally:
        b.close()


def test_autotuner_times_candidates_and_caches_the_choice(tmp_path):
    sizes = [(96, 64, 64), (64, 48, 128), (40, 30, 24)]
    specs, refs = make_specs(sizes, 4000)
    b1 = sgm.Batch(sgm.SgmParams(**BASE), specs)
    p1 = b1.plan()
    b1.close()
    # 20 line plans and 6 band-sweep plans (3 band heights x 2 tiles)
    assert p1["candidates_timed"] == 26 and p1["from_cache"] == 0 and p1["best_ms"] <= p1["worst_ms"]
    knobs = ("cost_volume", "path_lines", "branches", "path_group", "sweep_band", "sweep_tile")
    b2 = sgm.Batch(sgm.SgmParams(**BASE), specs)
    try:
        p2 = b2.plan()
        assert p2["from_cache"] == 1 and all(p2[k] == p1[k] for k in knobs)
        b2.run()
        b2.synchronize()
        for i in range(len(sizes)):
            assert_same(refs[i], b2.results(i), FIELDS, tag=f"tuned pair {i}")
    finally:
        b2.close()
    # A knob fixed against the cached choice is honoured: the cached plan is not taken, the open knobs
    # are searched, and the restricted result does not replace the cached plan.
    other = 1 - p1["cost_volume"]
    b3 = sgm.Batch(sgm.SgmParams(**BASE), specs, tuning={"cost_volume": other})
    try:
        p3 = b3.plan()
        assert p3["cost_volume"] == other and p3["from_cache"] == 0 and p3["candidates_timed"] > 1
        b3.run()
        b3.synchronize()
        for i in range(len(sizes)):
            assert_same(refs[i], b3.results(i), FIELDS, tag=f"restricted pair {i}")
    finally:
        b3.close()
    b4 = sgm.Batch(sgm.SgmParams(**BASE), specs)
    p4 = b4.plan()
    b4.close()
    assert p4["from_cache"] == 1 and all(p4[k] == p1[k] for k in knobs)
    # A knob fixed in agreement with the cached plan takes it from the cache.
    b5 = sgm.Batch(sgm.SgmParams(**BASE), specs, tuning={"cost_volume": p1["cost_volume"]})
    p5 = b5.plan()
    b5.close()
    assert p5["from_cache"] == 1 and all(p5[k] == p1[k] for k in knobs)
    # Persisted: the saved file holds the plan; a file loaded later replaces it for the same key.
    path = tmp_path / "plans.txt"
    assert sgm.tune_cache_save(str(path)) >= 1
    # The key: path count, census window, slant flag, then each pair's width and height in 64-pixel
    # buckets, disparities per lane (p: padded) and median flag, sorted.
    def dpl(nd):  # the library's lane layouts: 4, 8, 12, 16, 24 or 32 disparities per lane
        return next(d for d in (4, 8, 12, 16, 24, 32) if 16 * d >= nd)
    items = sorted(f"{-(-w // 64)}x{-(-h // 64)}d{dpl(nd)}{'p' if 16 * dpl(nd) != nd else ''}m" for w, h, nd in sizes)
    key = "v3:8:4x3:s1;" + ";".join(items)
    assert key + "\t" + " ".join(str(p1[k]) for k in knobs) in path.read_text().splitlines(), (key, path.read_text())
    alt = dict(cost_volume=1 - p1["cost_volume"], path_lines=4, branches=2, path_group=16, sweep_band=0, sweep_tile=64)
    if alt["cost_volume"] == 1:  # a loaded band-sweep plan
        alt.update(path_lines=8, branches=1, sweep_band=16, sweep_tile=32)
    path.write_text(key + "\t" + " ".join(str(alt[k]) for k in knobs) + "\n")
    assert sgm.tune_cache_load(str(path)) == 1
    b6 = sgm.Batch(sgm.SgmParams(**BASE), specs)
    try:
        p6 = b6.plan()
        assert p6["from_cache"] == 1 and all(p6[k] == alt[k] for k in knobs)
        b6.run()
        b6.synchronize()
        for i in range(len(sizes)):
            assert_same(refs[i], b6.results(i), FIELDS, tag=f"loaded-plan pair {i}")
    finally:
        b6.close()


def batch_sweep(n, seed):
    """Seeded random batches: 2..9 pairs of random sizes (1..300 x 1..120), disparity counts
    (1..512) and min_disparity each, random shared SGM parameters and a random fixed plan."""
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        rx, ry = int(rng.integers(0, 5)), int(rng.integers(0, 4))
        cells = (2 * rx + 1) * (2 * ry + 1)
        params = dict(census_radius_x=rx, census_radius_y=ry, p1=int(rng.integers(0, 255 - cells)),
                      p2=int(rng.integers(0, 256 - cells)), adaptive_p2=bool(rng.random() < 0.5),
                      eight_paths=bool(rng.random() < 0.6), slant_steps=bool(rng.random() < 0.5),
                      uniqueness=float(np.float32(rng.uniform(0.5, 1.0))), lr_max_diff=int(rng.integers(0, 3)),
                      subpixel=bool(rng.random() < 0.8))
        pairs = [(int(rng.integers(1, 301)), int(rng.integers(1, 121)),
                  int(rng.choice([1, 7, 16, 33, 64, 96, 128, 130, 192, 256, 384, 512])), int(rng.integers(-40, 41)))
                 for _ in range(int(rng.integers(2, 10)))]
        tuning = dict(cost_volume=int(rng.integers(0, 2)), path_lines=int(rng.choice([4, 8, 16])),
                      branches=int(rng.integers(0, 3)), path_group=int(rng.choice([16, 32])))
        if tuning["path_group"] == 32:
            tuning["path_lines"] = min(tuning["path_lines"], 8)  # 32-lane scanlines take at most 8 lines
        # Band sweeps (8 paths with the cost volume) in about half of those batches.
        sweep = params["eight_paths"] and tuning["cost_volume"] == 1 and rng.random() < 0.6
        tuning["sweep_band"] = int(rng.integers(1, 70)) if sweep else 0
        tuning["sweep_tile"] = int(rng.choice([16, 32, 48, 64]))
        cases.append((params, pairs, tuning, 5000 + 100 * i))
    return cases


BATCH_SWEEP = batch_sweep(8, 606)


@pytest.mark.parametrize("params,pairs,tuning,seed0", BATCH_SWEEP,
                         ids=[f"batch{i}-{len(c[1])}pairs" for i, c in enumerate(BATCH_SWEEP)])
def test_random_batches_bit_exact(params, pairs, tuning, seed0):
    specs, refs = [], []
    for i, (w, h, nd, min_d) in enumerate(pairs):
        pair = speckle_pair(w, h, max(nd, 6), seed0 + i)
        specs.append(sgm.PairSpec(pair.left, pair.right, min_d, nd, outputs=tuple(FIELDS)))
        refs.append(O.sgm(pair.left, pair.right, O.SgmParams(min_disparity=min_d, num_disparities=nd, **params)))
    print(f"random batch {pairs} {params} {tuning}", flush=True)
    b = sgm.Batch(sgm.SgmParams(**params), specs, tuning=tuning)
    try:
        for mode in ("graph", "plain", "graph"):
            b.clear_outputs(0xEE)
            b.run() if mode == "graph" else b.run_plain()
            b.synchronize()
            for i in range(len(pairs)):
                assert_same(refs[i], b.results(i), FIELDS, tag=f"{mode} pair {i} {pairs[i]}")
    finally:
        b.close()


@pytest.mark.parametrize("sweep_band", [0, 8], ids=["lines", "band-sweeps"])
def test_batch_with_wide_and_narrow_pairs(sweep_band):
    """A lane class holding a pair wider than 4096 columns runs the wide row kernel for all its pairs
    (over band-sweep records too); every pair equals the oracle, graph and plain."""
    sizes = [(5000, 20, 64), (300, 80, 64), (4100, 12, 128), (200, 50, 128)]
    specs, refs = make_specs(sizes, 6000)
    b = sgm.Batch(sgm.SgmParams(**BASE), specs, tuning=dict(cost_volume=1, path_lines=8, branches=1, path_group=16,
                                                            sweep_band=sweep_band, sweep_tile=64))
    try:
        for mode in ("graph", "plain"):
            b.clear_outputs(0xEE)
            b.run() if mode == "graph" else b.run_plain()
            b.synchronize()
            for i in range(len(sizes)):
                assert_same(refs[i], b.results(i), FIELDS, tag=f"wide batch {mode} pair {i} {sizes[i]}")
    finally:
        b.close()


def test_band_sweeps_on_degenerate_shapes():
    """Band sweeps on a 1x1 pair, a single column, a single row, three rows and a pair narrower than
    its disparity range, with bands of 2 rows and 16-column tiles: every pair equals the oracle."""
    sizes = [(1, 1, 16), (1, 40, 32), (40, 1, 64), (17, 3, 128), (9, 33, 96)]
    specs, refs = make_specs(sizes, 6500)
    b = sgm.Batch(sgm.SgmParams(**BASE), specs, tuning=dict(cost_volume=1, path_lines=4, branches=1, path_group=16,
                                                            sweep_band=2, sweep_tile=16))
    try:
        b.clear_outputs(0xEE)
        b.run()
        b.synchronize()
        for i in range(len(sizes)):
            assert_same(refs[i], b.results(i), FIELDS, tag=f"degenerate sweep pair {i} {sizes[i]}")
    finally:
        b.close()
