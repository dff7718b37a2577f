"""The C-ABI library loads and exports every function include/*.h declares; host-side argument
checks answer with einstar::Errc codes without touching a GPU."""
import ctypes as C
import glob
import os
import re

import pytest

from open_source_einstar_amd import sgm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(sgm_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("sgm_create", "sgm_compute", "sgm_compute_device", "sgm_destroy", "sgm_last_error",
              "sgm_last_timings", "sgm_batch_create", "sgm_batch_run"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = sgm.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_defaults_match_reference_sgm_params():
    p = sgm.CSgmParams()
    sgm.lib().sgm_params_default(C.byref(p))
    ref = sgm.SgmParams()  # stereo.hpp:26-40
    for f, _ in sgm.CSgmParams._fields_:
        assert getattr(p, f) == pytest.approx(float(getattr(ref, f)))


@pytest.mark.parametrize("bad", [
    dict(num_disparities=0), dict(num_disparities=513), dict(census_radius_x=5), dict(census_radius_y=4),
    dict(p1=-1), dict(p2=-1), dict(census_radius_x=4, census_radius_y=3, p2=200),
    dict(census_radius_x=4, census_radius_y=3, p1=(1 << 30) + 1, p2=10),
])
def test_invalid_parameters_are_rejected_before_any_gpu_work(bad):
    with pytest.raises(sgm.SgmError) as e:
        sgm.HipStereo.create(sgm.SgmParams(**bad), 64, 48)
    assert e.value.errc == "invalid_argument"


def test_penalty_limit_is_metal_stereos():
    """cells + P2 <= 255 is MetalStereo::create's limit and message (metal_stereo.cpp:118-120); P1 has
    no such limit there (large P1 with adaptive P2 runs with u16 path costs here)."""
    with pytest.raises(sgm.SgmError) as e:
        sgm.HipStereo.create(sgm.SgmParams(census_radius_x=4, census_radius_y=3, p2=193), 64, 48)
    assert e.value.errc == "invalid_argument" and "census window + P2 exceed the byte-sized SGM costs" in str(e.value)


def test_invalid_sizes_are_rejected():
    for (w, h) in [(0, 10), (10, 0), (32769, 10)]:
        with pytest.raises(sgm.SgmError) as e:
            sgm.HipStereo.create(sgm.SgmParams(), w, h)
        assert e.value.errc == "invalid_argument"


def test_stereo_frames_beyond_2_31_pixels_are_rejected_before_any_gpu_work():
    """The full-resolution frame of a compute_disparity context is indexed with 32-bit pixel numbers
    (speckle labels): 32768 x 65537 with two pyramid levels passes the SGM level's census limit and is
    refused for its full-resolution size."""
    with pytest.raises(sgm.SgmError) as e:
        sgm.HipStereo.create_stereo(sgm.SgmParams(num_disparities=16), sgm.StereoParams(pyramid_levels=2), 32768, 65537)
    assert e.value.errc == "invalid_argument" and "2^31 - 1 pixels" in str(e.value)


@pytest.mark.parametrize("bad", [dict(pyramid_levels=5), dict(refine_search_radius=5), dict(refine_search_radius=0),
                                 dict(detail_search=5.0), dict(detail_step=0.0), dict(refine_zncc_radius=8)])
def test_invalid_stereo_parameters_are_rejected_before_any_gpu_work(bad):
    """MetalStereo::create's limits on the refinement stages (metal_stereo.cpp:145-149)."""
    with pytest.raises(sgm.SgmError) as e:
        sgm.HipStereo.create_stereo(sgm.SgmParams(num_disparities=16), sgm.StereoParams(**bad), 320, 256)
    assert e.value.errc == "invalid_argument"


def test_stereo_defaults_match_reference_stereo_params():
    import oracle as O
    p = sgm.CSgmStereoParams()
    sgm.lib().sgm_stereo_params_default(C.byref(p))
    ref = O.StereoParams()  # stereo.hpp:42-89
    for f, _ in sgm._STEREO_FIELDS:
        assert getattr(p, f) == pytest.approx(float(getattr(ref, f))), f


def test_struct_sizes_match_the_header():
    """The ctypes mirrors have the C structs' sizes (a field added to one side only would shift every
    later field)."""
    import subprocess
    import tempfile
    names = ["sgm_params", "sgm_outputs", "sgm_timings", "sgm_stage_totals", "sgm_stereo_params", "sgm_request",
             "sgm_server_config", "sgm_completion", "sgm_server_stats", "sgm_pair", "sgm_batch_tuning",
             "sgm_batch_plan", "sgm_rectified_geometry", "sgm_point_params", "sgm_frame"]
    src = '#include <stdio.h>\n#include "sgm_hip.h"\nint main(void) {\n' + "".join(
        f'printf("%zu\\n", sizeof({n}));\n' for n in names) + "return 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "s.c"), os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        sizes = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    mirrors = [sgm.CSgmParams, sgm.CSgmOutputs, sgm.CSgmTimings, sgm.CSgmStageTotals, sgm.CSgmStereoParams,
               sgm.CSgmRequest, sgm.CSgmServerConfig, sgm.CSgmCompletion, sgm.CSgmServerStats, sgm.CSgmPair,
               sgm.CSgmBatchTuning, sgm.CSgmBatchPlan, sgm.CSgmRectifiedGeometry, sgm.CSgmPointParams, sgm.CSgmFrame]
    for n, m, sz in zip(names, mirrors, sizes):
        assert C.sizeof(m) == sz, (n, C.sizeof(m), sz)


def test_noalloc_entry_point_refuses_oversize_pairs_before_any_gpu_work():
    """sgm_compute_device_noalloc checks the pair against its static scratch before it touches the
    device (the pointers here are never dereferenced)."""
    L = sgm.lib()
    L.sgm_compute_device_noalloc.argtypes = [C.POINTER(sgm.CSgmParams), C.c_void_p, C.c_void_p, C.c_ssize_t, C.c_int,
                                             C.c_int, C.POINTER(sgm.CSgmOutputs), C.c_void_p]
    L.sgm_noalloc_scratch_bytes.restype = C.c_size_t
    assert L.sgm_noalloc_scratch_bytes() == 128 << 20
    cp = sgm.SgmParams(num_disparities=128, eight_paths=True, census_radius_x=4, census_radius_y=3, p1=10,
                       p2=120).to_c()
    rc = L.sgm_compute_device_noalloc(C.byref(cp), 16, 16, 1280, 1280, 1024, C.byref(sgm.CSgmOutputs()), None)
    assert sgm.ERRC[rc] == "invalid_argument" and b"scratch" in L.sgm_last_error(None)
    bad = sgm.SgmParams(num_disparities=600).to_c()
    rc = L.sgm_compute_device_noalloc(C.byref(bad), 16, 16, 64, 64, 48, C.byref(sgm.CSgmOutputs()), None)
    assert sgm.ERRC[rc] == "invalid_argument"


@pytest.mark.parametrize("tuning,eight,p1", [
    (dict(sweep_band=-2), True, 10), (dict(sweep_band=1025), True, 10), (dict(sweep_tile=24), True, 10),
    (dict(sweep_tile=80), True, 10), (dict(sweep_band=16), False, 10),       # band sweeps need 8 paths
    (dict(sweep_band=16, cost_volume=0), True, 10),                         # ... and the cost volume
    (dict(sweep_band=16), True, 200),                                       # ... and byte path costs
])
def test_batch_tuning_refusals_before_any_gpu_work(tuning, eight, p1):
    """sgm_batch_create_tuned refuses band-sweep knobs it cannot honour before it allocates anything
    (the pair pointers here are never dereferenced). P1 = 200 with adaptive P2: cells (63) +
    max(P2, P1 + 1) > 255, so the path costs are u16."""
    L = sgm.lib()
    pairs = (sgm.CSgmPair * 1)()
    pairs[0].left, pairs[0].right, pairs[0].pitch = 16, 16, 64
    pairs[0].width, pairs[0].height, pairs[0].min_disparity, pairs[0].num_disparities = 64, 48, 0, 32
    pairs[0].out.disparity = 16
    t = dict(cost_volume=-1, path_lines=-1, branches=-1, path_group=-1, sweep_band=-1, sweep_tile=-1)
    t.update(tuning)
    tn = sgm.CSgmBatchTuning(*[t[f] for f, _ in sgm.CSgmBatchTuning._fields_])
    cp = sgm.SgmParams(num_disparities=32, eight_paths=eight, census_radius_x=4, census_radius_y=3, p1=p1,
                       p2=120).to_c()
    h = C.c_void_p()
    rc = L.sgm_batch_create_tuned(C.byref(cp), 1, pairs, C.byref(tn), C.byref(h))
    assert sgm.ERRC[rc] == "invalid_argument", (tuning, rc)
    assert b"tuning" in L.sgm_last_error(None) and not h.value
