"""Served SGM: frames admitted continuously by free HBM give, frame by frame, the bytes they give
alone (sgm_compute) and the oracle's; a small arena forces queueing; oversize frames are refused."""
import time

import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import speckle_pair
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

BASE = dict(census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=True)
SERVE_FIELDS = ["disparity", "disparity_median", "disparity_int", "status", "right_disparity", "best_cost",
                "second_cost"]


def frames(n, seed0):
    out = []
    for i in range(n):
        w, h, nd = [(320, 256, 64), (200, 160, 48), (256, 200, 96)][i % 3]
        p = speckle_pair(w, h, nd, seed0 + i)
        out.append(p)
    return out


def alone(p):
    kw = dict(BASE, min_disparity=p.min_disparity, num_disparities=p.num_disparities)
    h, w = p.left.shape
    with sgm.HipStereo.create(sgm.SgmParams(**kw), w, h) as st:
        return st.compute(p.left, p.right, outputs=SERVE_FIELDS)


def drain(sv, n, timeout_s=60):
    """Completions with their outputs copied out of the server's pinned pool (which is then reused)."""
    got = []
    t0 = time.time()
    while len(got) < n and time.time() - t0 < timeout_s:
        for c in sv.poll(64, timeout_us=100000):
            outs = {f: a.copy() for f, a in (c["outputs"] or {}).items()}
            sv.release(c)
            c["outputs"] = outs
            got.append(c)
    return got


def test_served_results_equal_alone_and_oracle():
    fs = frames(9, 500)
    print(f"serving {len(fs)} frames", flush=True)
    with sgm.Server(sgm.SgmParams(**BASE), 320, 256, 96, slots=3) as sv:
        tickets = {}
        for i, p in enumerate(fs):  # all submitted at once: several on the device together
            t = sv.submit(p.left, p.right, p.min_disparity, p.num_disparities, outputs=SERVE_FIELDS, tag=i)
            tickets[t] = i
        got = drain(sv, len(fs))
        st = sv.stats()
    assert len(got) == len(fs) and all(c["status"] == 0 for c in got)
    assert st["completed"] == len(fs) and st["failed"] == 0 and st["queued"] == 0
    for c in got:
        p = fs[c["tag"]]
        assert_same(alone(p), c["outputs"], SERVE_FIELDS, tag=f"served frame {c['tag']} vs alone")
        kw = dict(BASE, min_disparity=p.min_disparity, num_disparities=p.num_disparities)
        assert_same(O.sgm(p.left, p.right, O.SgmParams(**kw)), c["outputs"], SERVE_FIELDS, tag=f"frame {c['tag']} vs oracle")
        assert c["t_submit_ns"] <= c["t_admit_ns"] <= c["t_done_ns"] and c["gpu_ms"] > 0


def test_small_arena_queues_and_admits_in_order():
    fs = frames(6, 600)
    # Room for about one 320x256x64 8-path frame (~42 MB of path costs) at a time.
    with sgm.Server(sgm.SgmParams(**BASE), 320, 256, 96, slots=4, hbm_budget_bytes=96 << 20) as sv:
        for i, p in enumerate(fs):
            sv.submit(p.left, p.right, p.min_disparity, p.num_disparities, outputs=SERVE_FIELDS, tag=i)
        got = drain(sv, len(fs))
    assert len(got) == len(fs)
    admits = sorted(got, key=lambda c: c["t_admit_ns"])
    assert [c["tag"] for c in admits] == sorted(c["tag"] for c in got)  # arrival order
    for c in got:
        p = fs[c["tag"]]
        kw = dict(BASE, min_disparity=p.min_disparity, num_disparities=p.num_disparities)
        assert_same(O.sgm(p.left, p.right, O.SgmParams(**kw)), c["outputs"], SERVE_FIELDS, tag=f"queued frame {c['tag']}")


def test_oversize_and_invalid_requests_are_refused():
    p = frames(1, 700)[0]
    with sgm.Server(sgm.SgmParams(**BASE), 320, 256, 96, slots=2, hbm_budget_bytes=16 << 20) as sv:
        with pytest.raises(sgm.SgmError) as e:
            sv.submit(p.left, p.right, p.min_disparity, p.num_disparities)
        assert e.value.errc == "busy"
        with pytest.raises(sgm.SgmError) as e:
            sv.submit(np.zeros((300, 400), np.uint8), np.zeros((300, 400), np.uint8), 0, 64)
        assert e.value.errc == "invalid_argument"
        assert sv.stats()["completed"] == 0


def test_raw_frames_rectified_with_resident_maps():
    """Raw frames of two map sets (full frame and crop size) mixed with rectified frames: each result
    equals compute_raw alone and the oracle on the oracle's rectified pair."""
    from open_source_einstar_amd.synth import rectification_maps

    sets = [(640, 512, 660, 530, 64), (320, 256, 300, 240, 48)]  # map w, h, raw w, h, D
    maps = [rectification_maps(mw, mh, rw, rh, 70 + i) for i, (mw, mh, rw, rh, _) in enumerate(sets)]
    raws = [speckle_pair(rw, rh, 16, 80 + i) for i, (_, _, rw, rh, _) in enumerate(sets * 2)]
    rect = frames(2, 900)
    pinned = sgm.PinnedPool()
    with sgm.Server(sgm.SgmParams(**BASE), 320, 256, 96, slots=3) as sv:
        ids = [sv.add_rectification(m.left_x, m.left_y, m.right_x, m.right_y, s[2], s[3]) for m, s in zip(maps, sets)]
        assert ids == [1, 2]
        jobs = []
        for i, p in enumerate(raws):
            k = i % 2
            left, right, borrow = p.left, p.right, i >= 2
            if borrow:  # the second half is read in place from pinned host memory
                left, right = pinned.take(p.left.shape, np.uint8), pinned.take(p.right.shape, np.uint8)
                left[:], right[:] = p.left, p.right
            jobs.append(("raw", k, p, sv.submit(left, right, -4, sets[k][4], outputs=SERVE_FIELDS, tag=len(jobs),
                                                rectification=ids[k], borrow=borrow)))
            q = rect[i % 2]
            jobs.append(("rect", None, q, sv.submit(q.left, q.right, q.min_disparity, q.num_disparities,
                                                    outputs=SERVE_FIELDS, tag=len(jobs))))
        got = {c["tag"]: c for c in drain(sv, len(jobs))}
        with pytest.raises(sgm.SgmError):  # raw frames of the wrong size for the set
            sv.submit(raws[1].left, raws[1].right, -4, 64, rectification=ids[0])
    pinned.free_all()
    assert len(got) == len(jobs) and all(c["status"] == 0 for c in got.values())
    for tag, (kind, k, p, _) in enumerate(jobs):
        c = got[tag]
        if kind == "rect":
            assert_same(alone(p), c["outputs"], SERVE_FIELDS, tag=f"rectified frame {tag}")
            continue
        mw, mh, rw, rh, nd = sets[k]
        m = maps[k]
        kw = dict(BASE, min_disparity=-4, num_disparities=nd)
        with sgm.HipStereo.create(sgm.SgmParams(**kw), mw // 2, mh // 2) as st:
            st.set_rectification(m.left_x, m.left_y, m.right_x, m.right_y, rw, rh)
            a = st.compute_raw(p.left, p.right, outputs=SERVE_FIELDS)
        assert_same(a, c["outputs"], SERVE_FIELDS, tag=f"raw frame {tag} vs compute_raw alone")
        rl = O.rectify_half(p.left, m.left_x, m.left_y)
        rr = O.rectify_half(p.right, m.right_x, m.right_y)
        assert_same(O.sgm(rl, rr, O.SgmParams(**kw)), c["outputs"], SERVE_FIELDS, tag=f"raw frame {tag} vs oracle")


def test_random_request_mix_bit_exact():
    """40 seeded random requests (sizes 1..320 x 1..256, D 1..96, min_disparity -20..39) submitted
    with random gaps while earlier ones run and complete, polled in between: every result equals
    the oracle's."""
    rng = np.random.default_rng(808)
    reqs = []
    for i in range(40):
        w, h = int(rng.integers(1, 321)), int(rng.integers(1, 257))
        nd = int(rng.choice([1, 8, 16, 33, 48, 64, 96]))
        p = speckle_pair(w, h, max(nd, 6), 900 + i)
        reqs.append((p.left, p.right, int(rng.integers(-20, 40)), nd))
    print(f"serving {len(reqs)} random requests", flush=True)
    got = []

    def collect(timeout_us):
        for c in sv.poll(64, timeout_us=timeout_us):
            outs = {f: a.copy() for f, a in (c["outputs"] or {}).items()}
            sv.release(c)
            c["outputs"] = outs
            got.append(c)

    with sgm.Server(sgm.SgmParams(**BASE), 320, 256, 96, slots=4, hbm_budget_bytes=192 << 20) as sv:
        for i, (l, r, md, nd) in enumerate(reqs):
            sv.submit(l, r, md, nd, outputs=SERVE_FIELDS, tag=i)
            if rng.random() < 0.3:
                time.sleep(float(rng.uniform(0, 0.002)))
            if rng.random() < 0.3:
                collect(0)
        t0 = time.time()
        while len(got) < len(reqs) and time.time() - t0 < 60:
            collect(100000)
        st = sv.stats()
    assert len(got) == len(reqs) and all(c["status"] == 0 for c in got)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert sorted(c["tag"] for c in got) == list(range(len(reqs)))
    for c in got:
        l, r, md, nd = reqs[c["tag"]]
        kw = dict(BASE, min_disparity=md, num_disparities=nd)
        assert_same(O.sgm(l, r, O.SgmParams(**kw)), c["outputs"], SERVE_FIELDS, tag=f"request {c['tag']} {l.shape} D={nd}")


@pytest.mark.parametrize("group", [0, 64])
def test_lane_widths_timed_or_fixed(group):
    """A shape whose lane layout allows 16, 32 and 64 lanes per path scanline (D = 256): with the width
    left open the first requests time each width twice before one is kept; with 64 fixed every request
    runs it. Every frame gives the bytes it gives alone and the oracle's."""
    fs = [speckle_pair(200, 160, 256, 900 + i) for i in range(8)]
    print(f"serving {len(fs)} frames of 200x160x256, path_group {group}", flush=True)
    with sgm.Server(sgm.SgmParams(**BASE), 256, 200, 256, slots=2, path_group=group) as sv:
        for i, p in enumerate(fs):
            sv.submit(p.left, p.right, p.min_disparity, p.num_disparities, outputs=SERVE_FIELDS, tag=i)
        got = drain(sv, len(fs))
        st = sv.stats()
    assert len(got) == len(fs) and all(c["status"] == 0 for c in got) and st["failed"] == 0
    for c in got:
        p = fs[c["tag"]]
        kw = dict(BASE, min_disparity=p.min_disparity, num_disparities=p.num_disparities)
        assert_same(alone(p), c["outputs"], SERVE_FIELDS, tag=f"frame {c['tag']} vs alone")
        assert_same(O.sgm(p.left, p.right, O.SgmParams(**kw)), c["outputs"], SERVE_FIELDS, tag=f"frame {c['tag']} vs oracle")
