"""The HIP library against the committed golden fixtures (tests/golden/golden.npz + manifest.json),
with no CPU computation in the comparison: SGM outputs and the digests of its census words, cost and
sum volumes, compute_disparity at the reference defaults and with one pyramid level, rectification
through resident maps followed by SGM, and point frames with their packed depth, all bit for bit."""
import numpy as np
import pytest

from open_source_einstar_amd import sgm
from golden_fixtures import MANIFEST, arrays, assert_bits, digest, entries

pytestmark = pytest.mark.gpu

SGM_FIELDS = ["disparity", "disparity_median", "disparity_int", "status", "right_disparity", "best_cost",
              "second_cost"]


@pytest.mark.parametrize("e", entries("sgm"))
def test_sgm_matches_golden(e):
    a = arrays("sgm", e)
    print(f"golden sgm {e['name']} {e['params']}", flush=True)
    with sgm.HipStereo.create(sgm.SgmParams(**e["params"]), e["width"], e["height"]) as st:
        got = st.compute(a["left"], a["right"], outputs=SGM_FIELDS + list(e["volumes"]))
    for f in SGM_FIELDS:
        assert_bits(a[f], got[f], f"{e['name']} {f}")
    for v, d in e["volumes"].items():
        assert digest(np.ascontiguousarray(got[v], dtype=np.dtype(d["dtype"]))) == d, (e["name"], v)


@pytest.mark.parametrize("e", entries("chain"))
def test_compute_disparity_matches_golden(e):
    a = arrays("chain", e)
    print(f"golden compute_disparity {e['name']} {e['params']} {e['stereo']}", flush=True)
    fl = np.repeat(np.repeat(a["left"], 2, 0), 2, 1)
    fr = np.repeat(np.repeat(a["right"], 2, 0), 2, 1)
    with sgm.HipStereo.create_stereo(sgm.SgmParams(**e["params"]), sgm.StereoParams(**e["stereo"]), e["width"],
                                     e["height"]) as ctx:
        got = ctx.compute_disparity(a["left"], a["right"], fl, fr)
    for f in ("disparity", "confidence"):
        assert_bits(a[f], got[f], f"{e['name']} {f}")


@pytest.mark.parametrize("e", entries("rectify"))
def test_rectify_then_sgm_matches_golden(e):
    a = arrays("rectify", e)
    print(f"golden rectify {e['name']}", flush=True)
    with sgm.HipStereo.create(sgm.SgmParams(**e["params"]), e["map_width"] // 2, e["map_height"] // 2) as st:
        st.set_rectification(a["left_x"], a["left_y"], a["right_x"], a["right_y"], e["raw_width"], e["raw_height"])
        got = st.compute_raw(a["raw_left"], a["raw_right"], outputs=SGM_FIELDS, rectified=True)
    for f in ["rect_left", "rect_right"] + SGM_FIELDS:
        assert_bits(a[f], got[f], f"{e['name']} {f}")


@pytest.mark.parametrize("e", entries("points"))
def test_points_and_packing_match_golden(e):
    a = arrays("points", e)
    print(f"golden points {e['name']}", flush=True)
    got = sgm.disparity_to_points(a["disparity"], a["confidence"], sgm.RectifiedGeometry(*e["geometry"]),
                                  sgm.PointParams(**e["point_params"]))
    assert_bits(a["points"], got["points"][..., :3], "points")
    assert_bits(a["normals"], got["normals"][..., :3], "normals")
    assert_bits(a["weights"], got["weights"], "weights")
    np.testing.assert_array_equal((got["points"][..., 3] > 0), a["points"][..., 2] > 0)
    assert_bits(a["packed"], sgm.pack_depth(got["points"], got["weights"]), "packed")


def test_every_family_is_covered():
    assert all(MANIFEST[f] for f in ("sgm", "chain", "rectify", "points"))
