"""The committed golden fixtures (tests/golden/golden.npz + manifest.json, written by
tests/golden/make_golden.py) against the oracle: the files are intact, the fixtures are not vacuous,
and the oracle reproduces every stored output and volume digest bit for bit on the stored inputs."""
import numpy as np
import pytest

import oracle as O
from golden_fixtures import DATA, MANIFEST, arrays, assert_bits, digest, entries


def test_files_match_the_manifest():
    n = 0
    for family, cases in MANIFEST.items():
        for e in cases:
            for k, d in e["arrays"].items():
                assert digest(DATA[f"{family}.{e['name']}.{k}"]) == d, (family, e["name"], k)
                n += 1
    assert n == len(DATA.files)


def test_fixtures_are_not_vacuous():
    for family in ("sgm", "chain", "rectify"):
        for e in MANIFEST[family]:
            assert (arrays(family, e)["disparity"] > -1e8).mean() > 0.5, (family, e["name"])
    for e in MANIFEST["points"]:
        assert (arrays("points", e)["points"][..., 2] > 0).mean() > 0.9, e["name"]


@pytest.mark.parametrize("e", entries("sgm"))
def test_oracle_sgm(e):
    a = arrays("sgm", e)
    ref = O.sgm(a["left"], a["right"], O.SgmParams(**e["params"]), volumes=True)
    for f in a:
        if f not in ("left", "right"):
            assert_bits(a[f], ref[f], f"{e['name']} {f}")
    for v, d in e["volumes"].items():
        assert digest(ref[v]) == d, (e["name"], v)


@pytest.mark.parametrize("e", entries("chain"))
def test_oracle_compute_disparity(e):
    a = arrays("chain", e)
    fl = np.repeat(np.repeat(a["left"], 2, 0), 2, 1)
    fr = np.repeat(np.repeat(a["right"], 2, 0), 2, 1)
    ref = O.compute_disparity(a["left"], a["right"], O.SgmParams(**e["params"]), O.StereoParams(**e["stereo"]),
                              full_left=fl, full_right=fr)
    for f in ("disparity", "confidence"):
        assert_bits(a[f], ref[f], f"{e['name']} {f}")


@pytest.mark.parametrize("e", entries("rectify"))
def test_oracle_rectify_then_sgm(e):
    a = arrays("rectify", e)
    rl = O.rectify_half(a["raw_left"], a["left_x"], a["left_y"])
    rr = O.rectify_half(a["raw_right"], a["right_x"], a["right_y"])
    assert_bits(a["rect_left"], rl, "rect_left")
    assert_bits(a["rect_right"], rr, "rect_right")
    ref = O.sgm(rl, rr, O.SgmParams(**e["params"]))
    for f in ("disparity", "disparity_median", "disparity_int", "status", "right_disparity", "best_cost",
              "second_cost"):
        assert_bits(a[f], ref[f], f"{e['name']} {f}")


@pytest.mark.parametrize("e", entries("points"))
def test_oracle_points_and_packing(e):
    a = arrays("points", e)
    ref = O.disparity_to_points(a["disparity"], a["confidence"], O.Geometry(*e["geometry"]),
                                O.PointParams(**e["point_params"]))
    for f in ("points", "normals", "weights"):
        assert_bits(a[f], ref[f], f"{e['name']} {f}")
    assert_bits(a["packed"], O.pack_depth(ref["points"], ref["weights"]), "packed")
