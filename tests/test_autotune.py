"""Tuning-record bookkeeping (no GPU): validity by plan source, bucket choice."""
import json
import os

from dogqc_amd import autotune


def _record(plan="c1_refill", block=256, rows=1, bucket=22, source=None):
    if source is None:
        source = autotune.planKey(autotune.translateVariant(plan, block, rows).source)
    return {"plan": plan, "block": block, "rows_per_thread": rows, "bucket": bucket, "plan_source": source,
            "module": "whatever.so", "tuning": {"grid": [512]}, "total_ms": 1.0}


def test_record_of_current_plan_source_is_current():
    assert autotune.isCurrent(_record())


def test_record_of_changed_plan_source_is_ignored():
    assert not autotune.isCurrent(_record(source="0" * 20))
    rec = _record()
    del rec["plan_source"]
    assert not autotune.isCurrent(rec)


def test_plan_key_differs_between_variants():
    a = autotune.planKey(autotune.translateVariant("c1_refill", 256, 1).source)
    b = autotune.planKey(autotune.translateVariant("c1_refill", 512, 1).source)
    c = autotune.planKey(autotune.translateVariant("c1", 256, 1).source)
    assert len({a, b, c}) == 3


def test_records_skip_stale_files(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE_DIR", str(tmp_path))
    autotune.saveRecord(_record(bucket=22))
    autotune.saveRecord(_record(bucket=19, source="f" * 20))
    recs = autotune.records("c1_refill")
    assert sorted(recs) == [22]
    with open(os.path.join(str(tmp_path), "c1_refill_n2e22.json")) as f:
        assert json.load(f)["bucket"] == 22


def test_nearest_bucket():
    recs = {19: "a", 22: "b", 26: "c"}
    assert autotune.nearest(recs, 1 << 22) == "b"
    assert autotune.nearest(recs, (1 << 20) + 5) == "a"      # bucket 20: 19 is 1 away, 22 is 2 away
    assert autotune.nearest(recs, 1 << 24) == "c"            # tie 22/26 -> the larger bucket
    assert autotune.nearest(recs, 1 << 29) is None           # 3 buckets from the nearest
    assert autotune.nearest({}, 100) is None
