"""Loads the committed golden fixtures (tests/golden/golden.npz + manifest.json, written by
tests/golden/make_golden.py) for the CPU and GPU golden tests."""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))
DATA = np.load(os.path.join(HERE, "golden.npz"))


def digest(a) -> dict:
    a = np.ascontiguousarray(a)
    return {"dtype": a.dtype.str, "shape": list(a.shape), "sha256": hashlib.sha256(a.tobytes()).hexdigest()}


def arrays(family, entry):
    return {k: DATA[f"{family}.{entry['name']}.{k}"] for k in entry["arrays"]}


def assert_bits(want, got, tag):
    want, got = np.asarray(want), np.ascontiguousarray(got, dtype=np.asarray(want).dtype)
    assert want.shape == got.shape, (tag, want.shape, got.shape)
    if not np.array_equal(want.view(np.uint8), got.view(np.uint8)):
        idx = np.argwhere(want != got)
        raise AssertionError(f"{tag}: {len(idx)} elements differ, first at {idx[0].tolist() if len(idx) else '?'}")


def entries(family):
    return [pytest.param(e, id=e["name"]) for e in MANIFEST[family]]
