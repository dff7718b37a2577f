"""The serving loop's admission logic on the CPU with a fake device (tests/cpp/serve_core_test.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "serve_core_test")


def test_admission_logic_with_fake_device():
    if not os.path.exists(EXE):
        from open_source_einstar_amd import build
        build.build_serve_core_test()
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "serve_core test ok" in r.stdout
