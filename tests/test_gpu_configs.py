"""The batched, served and large-batched workloads at their full sizes, each pair against the oracle:

* 32 uneven pairs (widths 320-1280, heights 240-1024, D in {64, 96, 128, 192}; the benchmark's batch)
  through the tuned graph and through the plain per-pair launches;
* a served stream of the 3:1 mix of 2560x2048 raw frames (rectified to 1280x1024, D = 128) and
  1280x1024 raw ROI readouts (rectified to 640x512, D = 64) through resident map sets, submitted with
  Poisson gaps at 60 Hz while earlier requests run: each result equals compute_raw alone and the
  oracle on the oracle's rectified pair;
* a resident batch of two 2048x1536 pairs with 256 disparities, streamed (inputs replaced, graph
  replayed) and launched plain;
* whole frames through the path-kernel scanline widths and row-kernel block sizes chosen per pair.
"""
import os

import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import rectification_maps, speckle_pair
from test_gpu_parity import FIELDS, assert_same

pytestmark = pytest.mark.gpu

PARAMS = dict(census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=True, adaptive_p2=True,
              slant_steps=True, uniqueness=0.9, lr_max_diff=1, subpixel=True)
THREADS = min(len(os.sched_getaffinity(0)), 64)


def uneven_sizes(n=32, seed=3):
    """The benchmark's batch: widths U[320, 1280], heights U[240, 1024], D from {64, 96, 128, 192}."""
    rng = np.random.default_rng(seed)
    ws, hs, ds = rng.integers(320, 1281, n), rng.integers(240, 1025, n), rng.choice([64, 96, 128, 192], n)
    return [(int(w), int(h), int(d)) for w, h, d in zip(ws, hs, ds)]


def test_batched_32_uneven_pairs():
    sizes = uneven_sizes()
    pairs = [speckle_pair(w, h, d, seed=1000 + i) for i, (w, h, d) in enumerate(sizes)]
    specs = [sgm.PairSpec(p.left, p.right, p.min_disparity, d, outputs=tuple(FIELDS)) for p, (_, _, d) in zip(pairs, sizes)]
    print(f"batch of {len(sizes)} pairs, {sum(w * h * d for w, h, d in sizes) / 1e6:.0f} Mpix*disp", flush=True)
    b = sgm.Batch(sgm.SgmParams(**PARAMS), specs)
    try:
        refs = [O.sgm(p.left, p.right, O.SgmParams(min_disparity=p.min_disparity, num_disparities=d, **PARAMS),
                      threads=THREADS) for p, (_, _, d) in zip(pairs, sizes)]
        for mode in ("graph", "plain"):
            b.clear_outputs(0xEE)
            b.run() if mode == "graph" else b.run_plain()
            b.synchronize()
            for i, sz in enumerate(sizes):
                assert_same(refs[i], b.results(i), FIELDS, tag=f"{mode} pair {i} {sz}")
        print(f"plan {b.plan()}", flush=True)
    finally:
        b.close()
    # The same batch under band sweeps of each band height the autotuner tries.
    for band in (16, 32, 64):
        b = sgm.Batch(sgm.SgmParams(**PARAMS), specs, tuning=dict(cost_volume=1, path_lines=8, branches=1,
                                                                   path_group=16, sweep_band=band, sweep_tile=64))
        try:
            b.clear_outputs(0xEE)
            b.run()
            b.synchronize()
            for i, sz in enumerate(sizes):
                assert_same(refs[i], b.results(i), FIELDS, tag=f"sweeps band {band} pair {i} {sz}")
            print(f"band sweeps of {band} rows bit-exact", flush=True)
        finally:
            b.close()


SERVE_FIELDS = ["disparity", "disparity_median", "disparity_int", "status", "right_disparity", "best_cost",
                "second_cost"]


@pytest.fixture(scope="module")
def served_inputs():
    """Two full-frame and two ROI raw pairs with their map sets, and each one's expected results
    (oracle on the oracle-rectified pair; compute_raw alone)."""
    kinds = [dict(raw=(2560, 2048), nd=128, seeds=(7200, 7201), map_seed=31),
             dict(raw=(1280, 1024), nd=64, seeds=(7300, 7301), map_seed=32)]
    out = []
    for k in kinds:
        rw, rh = k["raw"]
        m = rectification_maps(rw, rh, rw, rh, k["map_seed"])
        items = []
        for sd in k["seeds"]:
            p = speckle_pair(rw, rh, 2 * k["nd"], seed=sd)
            md = p.min_disparity // 2
            kw = dict(PARAMS, min_disparity=md, num_disparities=k["nd"])
            rl, rr = O.rectify_half(p.left, m.left_x, m.left_y), O.rectify_half(p.right, m.right_x, m.right_y)
            ref = O.sgm(rl, rr, O.SgmParams(**kw), threads=THREADS)
            with sgm.HipStereo.create(sgm.SgmParams(**kw), rw // 2, rh // 2) as st:
                st.set_rectification(m.left_x, m.left_y, m.right_x, m.right_y, rw, rh)
                alone = st.compute_raw(p.left, p.right, outputs=SERVE_FIELDS)
            items.append((p, md, ref, alone))
        out.append((k, m, items))
    return out


@pytest.mark.parametrize("compute_streams,path_group", [(0, 0), (-1, 32), (2, 16)],
                         ids=["shared-compute-stream-timed-lanes", "per-slot-streams-32-lanes", "two-compute-streams-16-lanes"])
def test_served_stream_of_full_frames_and_rois(served_inputs, compute_streams, path_group):
    import time

    rng = np.random.default_rng(7)
    n = 16
    kinds = np.array([0, 0, 0, 1] * (n // 4))
    rng.shuffle(kinds)
    gaps = rng.exponential(1.0 / 60.0, n)
    pool = sgm.PinnedPool()
    with sgm.Server(sgm.SgmParams(**PARAMS), 1280, 1024, 128, slots=4, compute_streams=compute_streams,
                    path_group=path_group) as sv:
        ids = []
        for k, m, items in served_inputs:
            rw, rh = k["raw"]
            ids.append(sv.add_rectification(m.left_x, m.left_y, m.right_x, m.right_y, rw, rh))
        pinned = {}
        for ki, (_, _, items) in enumerate(served_inputs):
            for j, (p, _, _, _) in enumerate(items):
                a, b = pool.take(p.left.shape, np.uint8), pool.take(p.right.shape, np.uint8)
                a[:], b[:] = p.left, p.right
                pinned[(ki, j)] = (a, b)
        got, jobs = {}, []

        def collect(timeout_us):
            for c in sv.poll(64, timeout_us=timeout_us):
                outs = {f: a.copy() for f, a in (c["outputs"] or {}).items()}
                sv.release(c)
                c["outputs"] = outs
                got[c["tag"]] = c

        t_next = time.perf_counter()
        for i in range(n):
            ki = int(kinds[i])
            j = i % 2
            k, _, items = served_inputs[ki]
            p, md, _, _ = items[j]
            t_next += gaps[i]
            while time.perf_counter() < t_next:
                collect(0)
                time.sleep(0.0002)
            a, b = pinned[(ki, j)] if i % 3 else (p.left, p.right)  # a third copied at submission
            sv.submit(a, b, md, k["nd"], outputs=SERVE_FIELDS, tag=i, rectification=ids[ki], borrow=bool(i % 3))
            jobs.append((ki, j))
            print(f"submitted {i} kind {ki}", flush=True)
        t0 = time.time()
        while len(got) < n and time.time() - t0 < 60:
            collect(100000)
        st = sv.stats()
    pool.free_all()
    assert len(got) == n and st["failed"] == 0 and all(c["status"] == 0 for c in got.values())
    overlapped = sum(1 for i in range(1, n) if got[i]["t_admit_ns"] < got[i - 1]["t_done_ns"])
    print(f"{overlapped} of {n - 1} requests admitted before the previous one completed", flush=True)
    for i, (ki, j) in enumerate(jobs):
        _, _, items = served_inputs[ki]
        _, _, ref, alone = items[j]
        assert_same(alone, got[i]["outputs"], SERVE_FIELDS, tag=f"request {i} kind {ki} vs compute_raw alone")
        assert_same(ref, got[i]["outputs"], SERVE_FIELDS, tag=f"request {i} kind {ki} vs oracle")


def test_large_pairs_streamed_through_a_resident_batch():
    W, H, D = 2048, 1536, 256
    src = [speckle_pair(W, H, D, seed=5000 + i) for i in range(4)]
    fields = ("disparity", "disparity_median", "disparity_int", "status", "right_disparity", "best_cost", "second_cost")
    specs = [sgm.PairSpec(p.left, p.right, p.min_disparity, D, outputs=fields) for p in src[:2]]
    print("resident batch of two 2048x1536x256 pairs", flush=True)
    b = sgm.Batch(sgm.SgmParams(**PARAMS), specs)
    try:
        refs = [O.sgm(p.left, p.right, O.SgmParams(min_disparity=p.min_disparity, num_disparities=D, **PARAMS),
                      threads=THREADS) for p in src]
        b.run()
        b.synchronize()
        for i in range(2):
            assert_same(refs[i], b.results(i), list(fields), tag=f"resident pair {i}")
        # Streamed: the next pairs' images replace the inputs (same min_disparity range: same plan).
        for i in range(2):
            assert src[2 + i].min_disparity == src[i].min_disparity
            b.set_inputs(i, src[2 + i].left, src[2 + i].right)
        for mode in ("graph", "plain"):
            b.clear_outputs(0xEE)
            b.run() if mode == "graph" else b.run_plain()
            b.synchronize()
            for i in range(2):
                assert_same(refs[2 + i], b.results(i), list(fields), tag=f"streamed {mode} pair {2 + i}")
    finally:
        b.close()


FULL_SIZE_FORMS = [
    # (w, h, nd, params, cost_volume, recompute_dirs): the path-kernel and row-kernel forms chosen per pair
    (1280, 1024, 512, dict(census_radius_x=4, census_radius_y=3, p1=10, p2=120), 1, 0),  # 64 lanes, 32 per lane
    (1280, 1024, 256, PARAMS, 0, 0),                                                      # 64 lanes, recomputed
    (1280, 1024, 128, PARAMS, 1, 4),                                                      # 32 lanes, half recomputed
    (960, 768, 256, dict(), 1, 0),                                                        # 32 lanes, 4 paths
    (320, 256, 247, dict(), 1, 0),  # the reference front end's stage: 32 lanes, 1024-thread row blocks
    (2048, 200, 256, PARAMS, 1, 0),                                                       # 512-thread row blocks
]


@pytest.mark.parametrize("w,h,nd,kw,cv,rd", FULL_SIZE_FORMS,
                         ids=[f"{c[0]}x{c[1]}x{c[2]}-cv{c[4]}-rd{c[5]}" for c in FULL_SIZE_FORMS])
def test_full_size_pairs_with_the_forms_chosen_per_pair(w, h, nd, kw, cv, rd):
    """Whole frames through the scanline widths and row-block sizes the library picks per pair, against
    the oracle, every output field."""
    pair = speckle_pair(w, h, min(nd, 256), seed=w + h + nd)
    params = dict(kw, min_disparity=pair.min_disparity, num_disparities=nd)
    print(f"full-size form {w}x{h}x{nd} cv={cv} rd={rd}", flush=True)
    ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), threads=THREADS)
    with sgm.HipStereo.create(sgm.SgmParams(**params), w, h) as st:
        st.set_option(sgm.OPT_COST_VOLUME, cv)
        st.set_option(sgm.OPT_RECOMPUTE_DIRS, rd)
        got = st.compute(pair.left, pair.right, outputs=FIELDS)
    assert_same(ref, got, FIELDS, tag=f"{w}x{h}x{nd} cv={cv} rd={rd}")
