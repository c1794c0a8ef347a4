"""CPU checks of the autotuner's table handling (no device)."""
import json
import os

import mestrado_amd.autotune as T


def test_buckets_by_median_length():
    assert T.bucket_of([100, 200, 300]) == "<=256"
    assert T.bucket_of([300, 400, 2000]) == "<=1024"
    assert T.bucket_of([2000, 3000]) == ">1024"
    assert T.arch_key("gfx950:sramecc+:xnack-") == "gfx950"


def test_save_and_load_round_trip(tmp_path, monkeypatch):
    monkeypatch.setattr(T, "TUNING_DIR", str(tmp_path))
    best = {"chunk": 16, "s6p2": 0, "wg5": 0, "ms": 1.5}
    p = T.save("gfx999:x", "<=1024", best, "w")
    t = T.load("gfx999")
    assert t["default"] == "<=1024"
    assert t["buckets"]["<=1024"]["chunk"] == 16
    T.save("gfx999", ">1024", dict(best, chunk=8), "w2")
    t = T.load("gfx999")
    assert t["default"] == "<=1024" and t["buckets"][">1024"]["chunk"] == 8
    assert json.load(open(p))["arch"] == "gfx999"


def test_committed_tables_are_well_formed():
    if not os.path.isdir(T.TUNING_DIR):
        return
    for name in os.listdir(T.TUNING_DIR):
        t = json.load(open(os.path.join(T.TUNING_DIR, name)))
        assert t["default"] in t["buckets"]
        for entry in t["buckets"].values():
            assert set(entry) - {"ms", "workload"} <= set(T.KEYS) and "chunk" in entry
