"""The autotuner's persisted cache on the CPU: records round-trip through the
JSON file, entry counts bucket to powers of two, and an untuned shape (or
no record) means one launch."""

from __future__ import annotations

from fhelium_amd import autotune


def test_cache_round_trip(tmp_path):
    path = tmp_path / "t.json"
    c = autotune.TuningCache(path)
    assert len(c) == 0 and c.get("k") is None
    c.put("k", {"schedule": [8, 2], "choice": "8 entries x 2 streams"})
    again = autotune.TuningCache(path)
    assert again.get("k")["schedule"] == [8, 2]
    assert autotune.schedule_from_record(again.get("k")) == autotune.NttSchedule(8, 2)
    assert autotune.schedule_from_record(None) == autotune.ONE_LAUNCH


def test_corrupt_or_old_cache_is_ignored(tmp_path):
    path = tmp_path / "t.json"
    path.write_text("{not json")
    assert len(autotune.TuningCache(path)) == 0
    path.write_text('{"version": 0, "entries": {"k": {}}}')
    assert len(autotune.TuningCache(path)) == 0


def test_entry_buckets():
    assert [autotune.entry_bucket(e) for e in (1, 2, 3, 7, 8, 100, 128, 129)] == [1, 2, 2, 4, 8, 64, 128, 128]
