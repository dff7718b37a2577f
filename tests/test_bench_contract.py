"""bench.py's contract on the CPU: one GPU only (any other --gpus exits 2 with a message before any
device work), and the committed counter summary it reads for `traffic` holds the kernels it reports,
with HBM bytes that match their fetch and write counters and that the algorithmic bytes bound."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2 and r.stdout == ""
    assert "multi-GPU runs are not supported" in r.stderr


def test_traffic_summary_matches_its_counters():
    with open(os.path.join(ROOT, "profiles", "sgm_1280x1024x128_traffic.json")) as f:
        d = json.load(f)
    k = d["kernels"]
    for name in ("paths_kernel<8, false, true, 16, true>", "rows_kernel<8, false, 8>", "cost8_kernel<8>"):
        assert name in k, name
        e = k[name]
        assert e["hbm_bytes_per_launch"] == e["fetch_bytes"] + e["write_bytes"] > 0
    # The reported path kernel moves its cost reads (1280 x 1024 x 128 bytes per direction, 8
    # directions) and path-cost writes (the same again), and not many times more.
    cost_plane = 1280 * 1024 * 128
    p = k["paths_kernel<8, false, true, 16, true>"]
    assert 0.95 * 8 * cost_plane <= p["write_bytes"] <= 1.05 * 8 * cost_plane
    assert 0.9 * 8 * cost_plane <= p["fetch_bytes"] <= 1.2 * 8 * cost_plane
    assert all(c.startswith("rocprofv3") or "rocprofv3" in c for c in d["commands"])
