"""Writes the golden fixtures of the stereo path: `golden.npz` (inputs and per-pixel outputs) and
`manifest.json` (parameters, shapes, dtypes and SHA-256 of every array, and of the census words,
cost and sum volumes, which are kept as digests only).

The reference holds no golden vectors, known-answer tests or fixtures for SGM, rectification, the
ZNCC refinement or the point frame (its stereo tests are statistical bars:
tests/depth/test_stereo.cpp:95-141, tests/depth/test_metal_stereo.cpp:96-121), and its CPU path
cannot be compiled in this image (libs/depth/src/stereo.cpp includes <tbb/parallel_for.h>). These
fixtures are therefore the outputs of the CPU oracle (oracle/sgm_oracle.c), whose functions are pinned
by closed forms, brute force and the reference's quality bars in tests/test_oracle_*.py. They freeze
those results as data: tests/test_golden_oracle.py requires the oracle to reproduce them, and
tests/test_gpu_golden.py requires the HIP library to reproduce them without running the oracle.

Run from the repository root after building the oracle:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

SGM_FIELDS = ["disparity", "disparity_median", "disparity_int", "status", "right_disparity", "best_cost",
              "second_cost"]
SGM_VOLUMES = ["census_left", "census_right", "cost_volume", "sum_volume"]
CHAIN_FIELDS = ["disparity", "confidence"]
POINT_FIELDS = ["points", "normals", "weights"]

# SGM on speckle pairs: both sizes, D in {16, 32}, 4 and 8 paths, adaptive P2 and slant steps on and
# off, min_disparity in {0, -4, 7}, census 5x5 (defaults), 9x7 and 3x3.
SGM_CASES = [
    ("64x48x16-4p", 64, 48, 16, 1, dict(min_disparity=0)),
    ("64x48x16-8p-fixed-p2", 64, 48, 16, 2, dict(min_disparity=-4, eight_paths=True, adaptive_p2=False)),
    ("64x48x32-8p-no-slant", 64, 48, 32, 3, dict(min_disparity=7, eight_paths=True, slant_steps=False)),
    ("96x64x32-4p-plain", 96, 64, 32, 4, dict(min_disparity=0, adaptive_p2=False, slant_steps=False)),
    ("96x64x16-8p-census3x3", 96, 64, 16, 5, dict(min_disparity=7, eight_paths=True, census_radius_x=1,
                                                   census_radius_y=1, uniqueness=0.8)),
    ("96x64x32-8p-census9x7", 96, 64, 32, 6, dict(min_disparity=-4, eight_paths=True, census_radius_x=4,
                                                   census_radius_y=3, p1=10, p2=120)),
    ("96x64x32-8p-no-subpixel", 96, 64, 32, 7, dict(min_disparity=0, eight_paths=True, subpixel=False,
                                                     lr_max_diff=0)),
]

# compute_disparity at the reference defaults (two pyramid levels, detail refinement on the
# full-resolution pair, occlusion and spike filters, speckle removal) and with one level.
CHAIN_SGM = dict(census_radius_x=4, census_radius_y=3, p1=10, p2=120, eight_paths=True)
CHAIN_CASES = [
    ("160x128-defaults", 160, 128, 48, 11, 2, 16, dict()),
    ("160x128-1-level-speckles-60", 160, 128, 48, 12, 1, 24, dict(pyramid_levels=1, speckle_max_region_size=60)),
]

# Rectification (remap at the maps' resolution, then downsample2) followed by SGM.
RECTIFY_CASES = [
    ("map64x48-raw70x52", 64, 48, 70, 52, 16, 21),
]

# Point frames: (f, cx, cy, baseline, cx_offset), point parameters.
POINT_CASES = [
    ("96x64-defaults", 96, 64, 31, (600.0, 47.5, 31.5, 50.0, 0.0), dict()),
    ("96x64-offset-depth-weight", 96, 64, 32, (580.25, 45.0, 30.75, 63.5, 12.5),
     dict(min_depth=100.0, max_depth=900.0, max_depth_jump=2.0, edge_weight=0.4, weight_reference_depth=320.0)),
]


def digest(a) -> dict:
    a = np.ascontiguousarray(a)
    return {"dtype": a.dtype.str, "shape": list(a.shape), "sha256": hashlib.sha256(a.tobytes()).hexdigest()}


def point_disparity(w, h, seed):
    """Two slanted planes with a depth jump, noise, holes and a too-near patch."""
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float32)
    d = (50 + 0.05 * xs - 0.03 * ys).astype(np.float32)
    d[:, w // 2:] += np.float32(6)
    d += rng.normal(0, 0.05, d.shape).astype(np.float32)
    d[rng.random(d.shape) < 0.05] = np.float32(-1e9)
    d[0:3, 0:5] = np.float32(2000)
    conf = rng.uniform(0.3, 1, (h, w)).astype(np.float32)
    return d, conf


def main():
    import oracle as O
    from open_source_einstar_amd.synth import rectification_maps, speckle_pair

    arrays, manifest = {}, {"sgm": [], "chain": [], "rectify": [], "points": []}

    def put(family, name, entry, values):
        entry["arrays"] = {}
        for k, v in values.items():
            key = f"{family}.{name}.{k}"
            arrays[key] = np.ascontiguousarray(v)
            entry["arrays"][k] = digest(v)
        manifest[family].append(entry)

    for name, w, h, nd, seed, kw in SGM_CASES:
        pair = speckle_pair(w, h, nd, seed)
        params = dict(num_disparities=nd, **kw)
        ref = O.sgm(pair.left, pair.right, O.SgmParams(**params), volumes=True)
        entry = dict(name=name, width=w, height=h, seed=seed, params=params,
                     volumes={k: digest(ref[k]) for k in SGM_VOLUMES})
        put("sgm", name, entry, dict(left=pair.left, right=pair.right, **{f: ref[f] for f in SGM_FIELDS}))

    for name, w, h, nd, seed, levels, nd_top, stereo in CHAIN_CASES:
        pair = speckle_pair(w, h, nd, seed, min_cells=2, max_cells=4)
        params = dict(min_disparity=pair.min_disparity // (1 << levels) - 1, num_disparities=nd_top, **CHAIN_SGM)
        full_l = np.repeat(np.repeat(pair.left, 2, 0), 2, 1)
        full_r = np.repeat(np.repeat(pair.right, 2, 0), 2, 1)
        ref = O.compute_disparity(pair.left, pair.right, O.SgmParams(**params), O.StereoParams(**stereo),
                                  full_left=full_l, full_right=full_r)
        entry = dict(name=name, width=w, height=h, seed=seed, params=params, stereo=stereo,
                     full_pair="each half-resolution pixel repeated 2x2")
        put("chain", name, entry, dict(left=pair.left, right=pair.right, **{f: ref[f] for f in CHAIN_FIELDS}))

    for name, mw, mh, rw, rh, nd, seed in RECTIFY_CASES:
        maps = rectification_maps(mw, mh, rw, rh, seed)
        raw = speckle_pair(rw, rh, 16, 100 + seed)
        params = dict(min_disparity=-4, num_disparities=nd, eight_paths=True, census_radius_x=4, census_radius_y=3,
                      p1=10, p2=120)
        rl = O.rectify_half(raw.left, maps.left_x, maps.left_y)
        rr = O.rectify_half(raw.right, maps.right_x, maps.right_y)
        ref = O.sgm(rl, rr, O.SgmParams(**params))
        entry = dict(name=name, map_width=mw, map_height=mh, raw_width=rw, raw_height=rh, seed=seed, params=params)
        put("rectify", name, entry, dict(raw_left=raw.left, raw_right=raw.right, left_x=maps.left_x,
                                         left_y=maps.left_y, right_x=maps.right_x, right_y=maps.right_y,
                                         rect_left=rl, rect_right=rr, **{f: ref[f] for f in SGM_FIELDS}))

    for name, w, h, seed, geom, pp in POINT_CASES:
        d, conf = point_disparity(w, h, seed)
        ref = O.disparity_to_points(d, conf, O.Geometry(*geom), O.PointParams(**pp))
        packed = O.pack_depth(ref["points"], ref["weights"])
        entry = dict(name=name, width=w, height=h, seed=seed, geometry=list(geom), point_params=pp)
        put("points", name, entry, dict(disparity=d, confidence=conf, packed=packed,
                                        **{f: ref[f] for f in POINT_FIELDS}))

    np.savez_compressed(os.path.join(HERE, "golden.npz"), **arrays)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"{len(arrays)} arrays, {os.path.getsize(os.path.join(HERE, 'golden.npz'))} bytes")


if __name__ == "__main__":
    main()
