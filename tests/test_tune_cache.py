"""Autotuner cache semantics on the CPU: an entry counts as a hit only when
it carries this layout version and every tuned dimension; older or partial
entries are misses (tuned again), never half-applied."""

from up3d_amd.tune.autotune import CACHE_VERSION, TUNED_KEYS, TuningCache


def _entry(**drop):
    e = {"version": CACHE_VERSION, "element_block": {"best": 256},
         "gather_chunk": {"best": 1024}, "walk_width": {"best": 2},
         "solver": {"best": {"restart": 200, "degree": 2, "max_coarse": 1000}}}
    for k in drop:
        e.pop(k)
    return e


def test_cache_hits_only_complete_current_entries(tmp_path):
    c = TuningCache(str(tmp_path / "t.json"))
    assert c.get("k") is None
    c.put("k", _entry())
    assert c.get("k")["walk_width"]["best"] == 2
    for k in TUNED_KEYS:
        c.put("k", _entry(**{k: None}))
        assert c.get("k") is None, k
    old = _entry()
    old.pop("version")                     # written before the cache was versioned
    c.put("k", old)
    assert c.get("k") is None
    c.put("k", dict(_entry(), version=CACHE_VERSION - 1))
    assert c.get("k") is None


def test_unreadable_cache_is_empty(tmp_path):
    p = tmp_path / "t.json"
    p.write_text("{not json")
    assert TuningCache(str(p)).get("k") is None
