"""Autotuner selection rule and cache (host logic; no GPU)."""
import json

from newrecon_amd import autotune as A
from newrecon_amd.batch import DEFAULT_TILE, TileCfg


class FakeBatch:
    def __init__(self, cases):
        self.cases = cases
        self.tile = TileCfg.from_dict(DEFAULT_TILE)

    def signature(self):
        return {"device": "gfx950", "cus": 256, "prj_dtype": "bfloat16", "potential": 2,
                "cases": self.cases}


def fake_measure(calls):
    def measure(kind, tile):
        calls.append((kind, tile.as_dict()))
        d = tile.as_dict()
        if kind == "fwd":  # fastest: rows 2, persistent 16 (two candidates tie)
            return 1.0 if d["fwd_rows"] == 2 and d["fwd_groups_per_cu"] in (16, 32) else 2.0
        return 0.5 if (d["upd_rows"], d["upd_waves"], d["upd_groups_per_cu"]) == (4, 4, 8) else 0.9
    return measure


def test_pick_takes_earliest_minimum():
    assert A.pick([3.0, 1.0, 1.0, 2.0]) == 1
    assert A.pick([1.0]) == 0


def test_tune_measures_then_caches(tmp_path):
    cache = str(tmp_path / "t.json")
    calls = []
    b = FakeBatch([[100, 50, 3, 256, 208, 3]])
    out = A.tune(b, cache_path=cache, measure=fake_measure(calls))
    assert not out["cached"]
    assert out["tile"]["fwd_rows"] == 2 and out["tile"]["fwd_groups_per_cu"] == 16
    assert (out["tile"]["upd_rows"], out["tile"]["upd_waves"],
            out["tile"]["upd_groups_per_cu"]) == (4, 4, 8)
    assert len(calls) == len(A.FWD_CANDIDATES) + len(A.UPD_CANDIDATES)
    # the update candidates are timed with the chosen forward shape in place
    assert all(c[1]["fwd_rows"] == 2 for c in calls if c[0] == "upd")
    assert b.tile.as_dict() == out["tile"]
    data = json.load(open(cache))
    assert list(data.values())[0]["tile"] == out["tile"]

    # a second tuner with the same cache measures nothing and agrees
    calls2 = []
    b2 = FakeBatch([[100, 50, 3, 256, 208, 3]])
    out2 = A.tune(b2, cache_path=cache, measure=fake_measure(calls2))
    assert out2["cached"] and calls2 == [] and out2["tile"] == out["tile"]
    assert b2.tile.as_dict() == out["tile"]

    # another signature is a cache miss
    calls3 = []
    A.tune(FakeBatch([[100, 50, 4, 256, 208, 3]]), cache_path=cache,
           measure=fake_measure(calls3))
    assert calls3 and len(json.load(open(cache))) == 2


def test_tune_without_cache_is_deterministic():
    outs = [A.tune(FakeBatch([[1, 2, 3, 4, 5, 6]]), cache_path=None,
                   measure=fake_measure([]))["tile"] for _ in range(3)]
    assert outs[0] == outs[1] == outs[2]


def test_signature_key_is_order_independent():
    a = {"b": 1, "a": [1, 2]}
    b = {"a": [1, 2], "b": 1}
    assert A.signature_key(a) == A.signature_key(b)
