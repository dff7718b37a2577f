"""Concurrency from host threads (the C calls release the interpreter lock, so the threads really
overlap): separate contexts used at the same time from different threads give each its own oracle
result, and one server fed by several submitting threads while another thread polls returns every
request's result bit-exact, each completion with its own outputs."""
import threading

import numpy as np
import pytest

import oracle as O
from open_source_einstar_amd import sgm
from open_source_einstar_amd.synth import speckle_pair
from test_gpu_parity import FIELDS, assert_same

pytestmark = pytest.mark.gpu

CONFIGS = [  # (w, h, nd, params)
    (320, 240, 64, dict(eight_paths=True, census_radius_x=4, census_radius_y=3, p1=10, p2=120)),
    (200, 150, 32, dict(eight_paths=False)),
    (257, 131, 96, dict(eight_paths=True, slant_steps=False, adaptive_p2=False, p1=7, p2=90)),
]


def test_contexts_used_concurrently_from_threads():
    cases = []
    for i, (w, h, nd, kw) in enumerate(CONFIGS):
        pairs = [speckle_pair(w, h, nd, seed=8100 + 10 * i + j) for j in range(3)]
        params = [dict(kw, min_disparity=p.min_disparity, num_disparities=nd) for p in pairs]
        refs = [O.sgm(p.left, p.right, O.SgmParams(**pp)) for p, pp in zip(pairs, params)]
        cases.append((w, h, pairs, params, refs))
    errors, results = [], {}

    def work(ci):
        w, h, pairs, params, refs = cases[ci]
        try:
            for rep in range(4):
                for j, (p, pp) in enumerate(zip(pairs, params)):
                    with sgm.HipStereo.create(sgm.SgmParams(**pp), w, h) as st:
                        results[(ci, rep, j)] = st.compute(p.left, p.right, outputs=FIELDS)
        except Exception as e:  # reported by the main thread
            errors.append((ci, repr(e)))

    threads = [threading.Thread(target=work, args=(ci,)) for ci in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads) and not errors, errors
    for (ci, rep, j), got in results.items():
        assert_same(cases[ci][4][j], got, FIELDS, tag=f"thread {ci} rep {rep} pair {j}")
    assert len(results) == 4 * sum(len(c[2]) for c in cases)


def test_server_fed_from_several_threads():
    w, h, nd = 320, 240, 64
    kw = dict(eight_paths=True, census_radius_x=4, census_radius_y=3, p1=10, p2=120)
    pairs = [speckle_pair(w, h, nd, seed=8200 + j) for j in range(4)]
    refs = [O.sgm(p.left, p.right, O.SgmParams(min_disparity=p.min_disparity, num_disparities=nd, **kw)) for p in pairs]
    fields = ["disparity", "disparity_median", "disparity_int", "status", "right_disparity", "best_cost", "second_cost"]
    n_threads, per_thread = 3, 8
    got, errors = {}, []
    with sgm.Server(sgm.SgmParams(**kw), w, h, nd, slots=4) as sv:
        done = threading.Event()

        def submitter(t):
            try:
                for k in range(per_thread):
                    j = (t + k) % len(pairs)
                    p = pairs[j]
                    sv.submit(p.left, p.right, p.min_disparity, nd, outputs=fields, tag=(t << 16) | (k << 4) | j)
            except Exception as e:
                errors.append(("submit", t, repr(e)))

        def poller():
            try:
                while len(got) < n_threads * per_thread and not done.is_set():
                    for c in sv.poll(16, timeout_us=20000):
                        outs = {f: a.copy() for f, a in (c["outputs"] or {}).items()}
                        sv.release(c)
                        c["outputs"] = outs
                        got[c["tag"]] = c
            except Exception as e:
                errors.append(("poll", repr(e)))

        pt = threading.Thread(target=poller)
        pt.start()
        ts = [threading.Thread(target=submitter, args=(t,)) for t in range(n_threads)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=60)
        pt.join(timeout=60)
        done.set()
        pt.join(timeout=10)
        st = sv.stats()
    assert not errors, errors
    assert len(got) == n_threads * per_thread and st["failed"] == 0
    for tag, c in got.items():
        assert c["status"] == 0 and c["outputs"], (tag, c["status"])
        assert_same(refs[tag & 0xF], c["outputs"], fields, tag=f"request {tag:#x}")
