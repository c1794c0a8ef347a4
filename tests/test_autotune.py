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
        if "served" in t:
            extra = {"ratio", "p99_ms", "offline_jobs_per_s", "workload"}
            assert set(t["served"]) - extra <= set(T.SERVE_KEYS)


def test_serve_keys_are_server_settings():
    # every setting the served search stores is one mst_server_set documents
    # (the call itself needs a device)
    here = os.path.dirname(os.path.abspath(__file__))
    head = open(os.path.join(here, "..", "include", "mestrado.h")).read()
    doc = head[:head.index("int mst_server_set(")]
    doc = doc[doc.rindex("/*"):]
    for k in T.SERVE_KEYS:
        assert '"%s"' % k in doc, k


def test_served_settings_round_trip(tmp_path, monkeypatch):
    monkeypatch.setattr(T, "TUNING_DIR", str(tmp_path))
    assert T.served_settings("gfx999") == []
    best = {"chunk": 12, "chunk6": 8, "s6p2": 1, "ratio": 2.5, "p99_ms": 200.0, "offline_jobs_per_s": 300.0}
    T.save_served("gfx999", best, "w")
    assert dict(T.served_settings("gfx999")) == {"chunk": 12, "chunk6": 8, "s6p2": 1}
    # a batched bucket saved afterwards keeps the served entry
    T.save("gfx999", "<=1024", {"chunk": 16, "ms": 1.0}, "w2")
    t = T.load("gfx999")
    assert t["served"]["chunk6"] == 8 and t["buckets"]["<=1024"]["chunk"] == 16
    assert list(T.serve_candidates())  # the search has candidates over SERVE_KEYS only
    assert all(set(c) == set(T.SERVE_KEYS) for c in T.serve_candidates())


class _FakeCtx:
    """Records the tunables the table applies (no device)."""

    def __init__(self, arch):
        self._arch, self.set_calls = arch, []

    def arch(self):
        return self._arch

    def _tune_set(self, key, value):
        self.set_calls.append((key, value))


def test_bucket_of_set_is_applied(tmp_path, monkeypatch):
    monkeypatch.setattr(T, "TUNING_DIR", str(tmp_path))
    T.save("gfx999", "<=1024", {"chunk": 64, "stack": 4, "ms": 1.0}, "w", default=True)
    T.save("gfx999", ">1024", {"chunk": 48, "stack": 3, "ms": 2.0}, "w")
    ctx = _FakeCtx("gfx999")
    applied = T.apply_for(ctx, [5000, 9000, 20000])
    assert applied["bucket"] == ">1024" and dict(ctx.set_calls) == {"chunk": 48, "stack": 3}
    ctx.set_calls.clear()
    assert T.apply_for(ctx, [100, 150]) == {}  # no <=256 bucket: nothing applied
    assert T.apply(ctx)["bucket"] == "<=1024" and dict(ctx.set_calls)["chunk"] == 64
