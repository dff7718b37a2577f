"""Pins of the oracle's disparity_to_points (libs/depth/src/point_image.cpp:10-55) and of the
session file's depth packing (libs/session/src/session.cpp:143-185): pinhole closed forms, the
analytic normal of planes, the view and depth-noise weights, the validity rules, and packing known
answers and round trips."""
import numpy as np
import pytest

import oracle as O

G = O.Geometry(f=600.0, cx=31.5, cy=23.25, baseline=50.0, cx_offset=4.0)


def test_points_of_a_constant_disparity_plane():
    h, w = 48, 64
    d = np.full((h, w), 56.0, np.float32)  # z = 600 * 50 / 60 = 500 mm
    o = O.disparity_to_points(d, None, G)
    z = 600.0 * 50.0 / (56.0 + 4.0)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    np.testing.assert_allclose(o["points"][..., 2], z, rtol=1e-7)
    np.testing.assert_allclose(o["points"][..., 0], (xs - G.cx) * z / G.f, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(o["points"][..., 1], (ys - G.cy) * z / G.f, rtol=1e-6, atol=1e-6)
    # Interior normals face the camera; the view weight is the worse of the two cameras' cosines.
    n = o["normals"][1:-1, 1:-1]
    np.testing.assert_allclose(n, np.broadcast_to([0.0, 0.0, -1.0], n.shape), atol=1e-6)
    p = o["points"].astype(np.float64)
    c1 = p[..., 2] / np.linalg.norm(p, axis=-1)
    q = p - np.array([G.baseline, 0, 0])
    c2 = q[..., 2] / np.linalg.norm(q, axis=-1)
    np.testing.assert_allclose(o["weights"][1:-1, 1:-1], np.minimum(c1, c2)[1:-1, 1:-1], rtol=1e-5)
    # Image borders have no normal: the edge weight.
    assert not o["normals"][0].any() and (o["weights"][0] == np.float32(0.25)).all()


def test_normals_of_a_tilted_plane():
    """Plane n.P = c: 1/Z is affine in the pixel, so the disparity is affine too; the central
    differences of the back-projected points are exact tangents, and the normal is the plane's."""
    h, w = 40, 50
    nrm = np.array([0.3, -0.2, -1.0])
    nrm /= np.linalg.norm(nrm)
    c = -450.0  # n.P = c with n facing the camera (z < 0 component), points near z = 450
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    ray = np.stack([(xs - G.cx) / G.f, (ys - G.cy) / G.f, np.ones_like(xs)], -1)
    Z = c / (ray @ nrm)
    d = (G.f * G.baseline / Z - G.cx_offset).astype(np.float32)
    o = O.disparity_to_points(d, np.full((h, w), 0.8, np.float32), G)
    got = o["normals"][1:-1, 1:-1].reshape(-1, 3)
    assert np.abs(got - nrm).max() < 2e-3
    assert (o["weights"][1:-1, 1:-1] > 0).all() and (o["weights"] <= 0.8).all()


def test_validity_rules():
    h, w = 10, 12
    d = np.full((h, w), 56.0, np.float32)
    d[2, 3] = -1.0e9                     # invalid disparity
    d[4, 4] = -4.0                       # d + cx_offset <= 0
    d[5, 5] = 600 * 50 / 140.0 - 4.0    # z = 140 < min_depth
    d[6, 6] = 600 * 50 / 800.0 - 4.0    # z = 800 > max_depth
    d[7, 7] = 600 * 50 / 700.0 - 4.0    # z = 700 == max_depth: kept
    o = O.disparity_to_points(d, None, G)
    for (y, x) in ((2, 3), (4, 4), (5, 5), (6, 6)):
        assert not o["points"][y, x].any() and o["weights"][y, x] == 0, (y, x)
    assert o["points"][7, 7, 2] > 0
    # Neighbours of a hole or across a depth jump (> 4 mm between the two neighbours) have no normal.
    assert not o["normals"][2, 2].any() and o["weights"][2, 2] == np.float32(0.25)
    assert not o["normals"][7, 6].any() and not o["normals"][6, 7].any()
    assert o["normals"][8, 9].any()


@pytest.mark.parametrize("z,factor", [(600.0, 0.25), (200.0, 1.0), (1000.0, 0.1), (300.0, 1.0)])
def test_depth_noise_weight(z, factor):
    p = O.PointParams(min_depth=10, max_depth=2000, weight_reference_depth=300.0)
    d = np.full((9, 9), np.float32(600 * 50 / z - 4.0), np.float32)
    a = O.disparity_to_points(d, None, G, p)
    b = O.disparity_to_points(d, None, G, O.PointParams(min_depth=10, max_depth=2000))
    np.testing.assert_allclose(a["weights"], b["weights"] * np.float32(factor), rtol=1e-6)


def test_pack_known_answers_and_round_trip():
    z = np.array([[0.0, 0.01, 0.02, 1.25, 1310.7, 2000.0, -5.0, 0.0104, 1.01]], np.float32)
    pts = np.zeros((1, z.shape[1], 3), np.float32)
    pts[..., 2] = z
    wt = np.array([[0.5 / 255, 1.2, -0.1, 0.5, 1.0, 0.0, 0.3, 0.002, 0.998]], np.float32)
    b = O.pack_depth(pts, wt)
    n = z.size
    q = np.frombuffer(b[:2 * n].tobytes(), np.uint16)
    absq = np.cumsum(q.astype(np.int64)) % 65536
    # lround(z * 50) in [1, 65535]: 1.25 -> 62.5 -> 63 (half away from zero); 0.0104 -> 0.52 -> 1;
    # 1.01f is 1.00999999 -> 50.4999995 -> 50 (the product is formed in double from the float)
    want = [0, 1, 1, 63, 65535, 65535, 0, 1, 50]
    assert absq.tolist() == want
    assert b[2 * n:].tolist() == [1, 255, 0, 128, 255, 0, 77, 1, 254]  # lround(w * 255) in [0, 255]
    depth, conf = O.unpack_depth(b, n, 1)
    np.testing.assert_allclose(depth[0], np.array(want) / 50.0, rtol=1e-7)
    # Random rows, including decreasing depth (deltas wrap modulo 2^16), decode to the quantised depth.
    rng = np.random.default_rng(1)
    pts = np.zeros((20, 31, 3), np.float32)
    pts[..., 2] = rng.uniform(-10, 1400, (20, 31))
    pts[rng.random((20, 31)) < 0.2, 2] = 0
    wt = rng.uniform(0, 1, (20, 31)).astype(np.float32)
    depth, conf = O.unpack_depth(O.pack_depth(pts, wt), 31, 20)
    zq = np.where(pts[..., 2] > 0, np.clip(np.round(pts[..., 2].astype(np.float64) * 50), 1, 65535) / 50, 0)
    np.testing.assert_allclose(depth, zq, atol=1e-4)
    assert np.abs(conf - wt).max() <= 0.5 / 255 + 1e-7
