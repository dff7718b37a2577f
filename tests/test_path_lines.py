"""The path kernel's scanline enumeration walks every pixel of every direction exactly once, as
whole scanlines (tests/cpp/path_lines_test.cpp, built from the kernel's own header)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "path_lines_test")


def test_every_pixel_walked_once_as_whole_scanlines():
    if not os.path.exists(EXE):
        from open_source_einstar_amd import build
        build.build_path_lines_test()
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
