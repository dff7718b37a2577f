"""The autotuner's persisted plan cache (sgm_tune_cache_load / _save), on the CPU: the file format
round-trips, and malformed or missing files are refused without touching the cache."""
import pytest

from open_source_einstar_amd import sgm


def test_round_trip(tmp_path):
    src = tmp_path / "plans.txt"
    key = "v3:8:4x3:s1;20x16d8m;5x4d4"
    src.write_text(f"{key}\t1 16 1 16 0 64\nv3:4:2x2:s0;1x1d2p\t0 4 2 32 0 64\nv3:8:4x3:s1;9x9d8\t1 8 1 16 32 48\n")
    assert sgm.tune_cache_load(str(src)) == 3
    out = tmp_path / "saved.txt"
    n = sgm.tune_cache_save(str(out))
    assert n >= 3
    lines = out.read_text().splitlines()
    assert f"{key}\t1 16 1 16 0 64" in lines and "v3:4:2x2:s0;1x1d2p\t0 4 2 32 0 64" in lines
    assert "v3:8:4x3:s1;9x9d8\t1 8 1 16 32 48" in lines
    # A later load replaces an entry of the same key.
    src.write_text(f"{key}\t0 8 2 32 0 64\n")
    assert sgm.tune_cache_load(str(src)) == 1
    sgm.tune_cache_save(str(out))
    assert f"{key}\t0 8 2 32 0 64" in out.read_text().splitlines()


@pytest.mark.parametrize("bad", ["no tab here\n", "k\t1 16 1 16\n", "k\t2 16 1 16 0 64\n", "k\t1 12 1 16 0 64\n",
                                 "k\t1 16 3 16 0 64\n", "k\t1 16 1 24 0 64\n", "\t1 16 1 16 0 64\n",
                                 "k\t0 8 1 16 32 64\n", "k\t1 8 1 16 -1 64\n", "k\t1 8 1 16 32 24\n",
                                 "k\t1 8 1 16 32 80\n"])
def test_malformed_lines_are_refused(tmp_path, bad):
    """Each refused line breaks one rule: a tab and six numbers, cost volume 0/1, 4/8/16 lines, 0-2
    branches, 16/32 lanes, band rows >= 0 and only with the cost volume, tiles of 16-64 columns in
    steps of 16."""
    f = tmp_path / "bad.txt"
    f.write_text("good\t1 16 1 16 0 64\n" + bad)
    with pytest.raises(sgm.SgmError) as e:
        sgm.tune_cache_load(str(f))
    assert e.value.errc == "protocol"
    out = tmp_path / "after.txt"
    sgm.tune_cache_save(str(out))
    assert "good\t1 16 1 16 0 64" not in out.read_text().splitlines()  # nothing of a refused file is merged


def test_missing_file(tmp_path):
    with pytest.raises(sgm.SgmError) as e:
        sgm.tune_cache_load(str(tmp_path / "absent.txt"))
    assert e.value.errc == "not_found"
