"""The manual build line in INTEGRATION.md names every source the build script compiles (a line that
misses one would not link)."""
import os
import re

from open_source_einstar_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_integration_build_line_lists_every_hip_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    listed = set(re.findall(r"\$S/([a-z0-9_]+\.hip)", text))
    assert listed == set(build.HIP_SOURCES), (sorted(listed), sorted(build.HIP_SOURCES))
    for s in build.HIP_SOURCES:
        assert os.path.exists(os.path.join(build.CSRC, s)), s
