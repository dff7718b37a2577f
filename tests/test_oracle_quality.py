"""The reference's own quality bars for SGM alone, applied to the oracle on synthetic speckle.

tests/depth/test_stereo.cpp:121-141 requires > 97% of valid SGM disparities within 0.5 px of the
truth on a fronto-parallel plane (its ray-cast renderer needs Eigen; the same scene comes from this
package's generator: one planar patch, no slope)."""
import numpy as np

import oracle as O
from open_source_einstar_amd.synth import speckle_pair


def test_fronto_parallel_plane_within_half_pixel():
    p = speckle_pair(320, 256, 64, seed=11, min_cells=1, max_cells=1, max_gradient=0.0)
    sp = O.SgmParams(min_disparity=p.min_disparity, num_disparities=64)
    d = O.sgm(p.left, p.right, sp)["disparity"]
    ys, xs = np.mgrid[20:236, 0:300]
    ys, xs = ys[xs > 120], xs[xs > 120]  # away from the unmatched left border
    v = d[ys, xs] > -1e8
    assert v.sum() > 0
    err = np.abs(d[ys, xs][v] - p.gt_disparity[ys, xs][v])
    assert (err < 0.5).mean() > 0.97


def test_piecewise_planar_scene_is_mostly_right():
    p = speckle_pair(320, 256, 64, seed=12)
    sp = O.SgmParams(min_disparity=p.min_disparity, num_disparities=64, census_radius_x=4, census_radius_y=3,
                     p1=10, p2=120, eight_paths=True)
    o = O.sgm(p.left, p.right, sp)
    d = o["disparity"]
    v = d > -1e8
    assert v.mean() > 0.6
    err = np.abs(d[v] - p.gt_disparity[v])
    assert np.median(err) < 0.35 and (err < 1.0).mean() > 0.9
    # median3 never makes a valid pixel invalid or vice versa.
    assert ((o["disparity_median"] > -1e8) == v).all()
