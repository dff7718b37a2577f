"""Served SGM: frames admitted continuously by free HBM give, frame by frame, the bytes they give
alone (sgm_compute) and the oracle's; a small arena forces queueing; oversize frames are refused."""
import time

import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import speckle_pair
from test_gpu_parity import FIELDS, assert_same

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
