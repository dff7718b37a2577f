"""The C++ drop-in class (include/einstar_hip/hip_stereo.hpp) on the GPU: its self-test program
checks create/compute/compute_sgm/last_timings and the error codes against the CPU oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "hip_stereo_selftest")


@pytest.mark.gpu
def test_hip_stereo_cpp_selftest():
    assert os.path.exists(EXE), "build first: python -m open_source_einstar_amd.build"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0 and "selftest ok" in r.stdout
