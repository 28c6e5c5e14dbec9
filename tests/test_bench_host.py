"""Host-side logic of bench.py (no GPU): the counter readings that the bench line quotes are used only
when they describe the build being timed."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _write(tmp_path, **fields):
    p = tmp_path / "counters.json"
    p.write_text(json.dumps(fields))
    return str(p)


def test_counters_fresh_only_for_the_profiled_sources_and_time(tmp_path, monkeypatch):
    h = bench.kernel_source_hash()
    monkeypatch.setattr(bench, "PAIR_RMU_COUNTERS", _write(tmp_path, kernel_ms=100.0, source_sha256=h))
    assert not bench.pair_counters(103.0)["stale"]
    assert bench.pair_counters(106.0)["stale"]            # another build's time
    monkeypatch.setattr(bench, "PAIR_RMU_COUNTERS", _write(tmp_path, kernel_ms=100.0, source_sha256="0" * 64))
    assert bench.pair_counters(100.0)["stale"]            # another build's sources
    monkeypatch.setattr(bench, "PAIR_RMU_COUNTERS", _write(tmp_path, kernel_ms=100.0))
    assert bench.pair_counters(100.0)["stale"]            # no hash recorded


def test_source_hash_covers_the_kernel_sources():
    for p in bench.PAIR_KERNEL_SOURCES:
        assert os.path.isfile(os.path.join(ROOT, p)), p
    assert len(bench.kernel_source_hash()) == 64


def test_committed_profile_is_readable():
    c = bench.pair_counters(1.0)
    assert c is not None
    for k in ("kernel_ms", "FETCH_SIZE_KiB", "FETCH_SIZE_scale", "WRITE_SIZE_KiB", "WRITE_SIZE_scale",
              "SQ_INSTS_VALU", "commands"):
        assert k in c
