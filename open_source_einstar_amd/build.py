"""Build recipe for the native parts, in-tree (no JIT caches): run `python -m open_source_einstar_amd.build`.

* ``libsgm_hip.so``   -- the gfx950 kernels and the C ABI of include/sgm_hip.h (hipcc)
* ``libeinstar_synth.so`` -- the synthetic speckle pair generator (gcc)
* ``hip_stereo_selftest`` -- C++ program driving the HipStereo wrapper (tests only)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIP_LIB = os.path.join(PKG, "libsgm_hip.so")
SYNTH_LIB = os.path.join(PKG, "libeinstar_synth.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: float stages (subpixel, uniqueness, left/right check) round like the reference.
# -packed-fp32-ops off: no hand-free v_pk_*_f32 (the kernels are integer; the flag guards the float bits).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wextra", "-Wno-unused-parameter",
]

# The path kernel's walk loop runs faster with its header aligned to 64 bytes (1280x1024x128, 8 paths:
# 459-460 against 468 us; 320x256x128: 47.8 against 53 us), where the code before it otherwise leaves it
# at any offset; the ZNCC kernels run about 1% slower aligned and the row kernel the same, so only the
# path-kernel sources take it.
PATH_SOURCE_FLAGS = ["-falign-loops=64"]

HIP_SOURCES = ["sgm_kernels.hip", "paths_g16_lo.hip", "paths_g16_hi.hip", "paths_g32.hip", "paths_g64.hip", "paths_w16.hip",
               "rows_kernels.hip", "rows_w16.hip", "sweep_kernels.hip", "rows_sweep.hip",
               "refine_kernels.hip", "frame_kernels.hip", "sgm_api.hip", "sgm_server.hip", "sgm_noalloc.hip"]
OBJ_DIR = os.path.join(PKG, "_build")


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip(verbose: bool = True) -> str:
    """Compiles each source to an object in parallel (the template-heavy kernel files dominate),
    then links the shared library."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES]
    jobs = [[HIPCC, *HIP_FLAGS, *(PATH_SOURCE_FLAGS if s.startswith("paths_") else []), f"-I{INCLUDE}", f"-I{CSRC}", "-c",
             os.path.join(CSRC, s), "-o", o]
            for s, o in zip(HIP_SOURCES, objs)]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1))
    with ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_synth() -> str:
    tmp = SYNTH_LIB + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off", "-Wall", "-Wextra", "-Werror",
          "-o", tmp, os.path.join(CSRC, "speckle_synth.c"), "-lm"])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


def build_oracle() -> str:
    _run(["make", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle_sgm.so")


def build_selftest() -> str | None:
    src = os.path.join(ROOT, "tests", "cpp", "hip_stereo_selftest.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "hip_stereo_selftest")
    # Links the built libraries by relative rpath so the binary also runs from a copy of the tree.
    _run(["g++", "-O2", "-std=c++20", "-Wall", "-Wextra", "-Werror", f"-I{INCLUDE}", f"-I{os.path.join(ROOT, 'oracle')}",
          "-o", out, src, f"-L{PKG}", "-lsgm_hip", f"-L{os.path.join(ROOT, 'oracle')}", "-loracle_sgm",
          "-Wl,-rpath,$ORIGIN/../../open_source_einstar_amd:$ORIGIN/../../oracle", "-lpthread", "-lm"])
    return out


def build_serve_core_test() -> str:
    out = os.path.join(ROOT, "tests", "cpp", "serve_core_test")
    _run(["g++", "-O2", "-std=c++20", "-Wall", "-Wextra", "-Werror", f"-I{CSRC}", "-o", out,
          os.path.join(ROOT, "tests", "cpp", "serve_core_test.cpp")])
    return out


def build_path_lines_test() -> str:
    out = os.path.join(ROOT, "tests", "cpp", "path_lines_test")
    _run(["g++", "-O2", "-std=c++20", "-Wall", "-Wextra", "-Werror", f"-I{CSRC}", "-o", out,
          os.path.join(ROOT, "tests", "cpp", "path_lines_test.cpp")])
    return out


def build_all() -> None:
    build_oracle()
    build_synth()
    build_hip()
    build_selftest()
    build_serve_core_test()
    build_path_lines_test()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "synth": build_synth, "oracle": build_oracle,
         "selftest": build_selftest, "serve_core": build_serve_core_test,
         "path_lines": build_path_lines_test}[w]()
