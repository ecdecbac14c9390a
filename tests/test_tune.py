"""Autotuner bookkeeping (CPU): knob set/get through the C ABI, the size
classes, nearest-entry choice and the JSON cache."""
import math

import pytest

import phem_native as P
import phem_tune as T


def test_set_get_tuning_round_trip():
    old = P.get_tuning()
    try:
        P.set_tuning(blocks_per_cu=8, mst_light_factor=0.5, cc_sample_factor=3.0)
        t = P.get_tuning()
        assert (t["blocks_per_cu"], t["mst_light_factor"], t["cc_sample_factor"]) == (8, 0.5, 3.0)
        # built-in kernel shapes
        assert {k: t[k] for k in ("mst_block", "mst_ept", "mst_ept_compact",
                                  "mst_compact_log2", "mst_split_ratio", "cc_block",
                                  "cc_ept", "cc_ept_filter", "cc_block_filter")} == {
            "mst_block": 512, "mst_ept": 4, "mst_ept_compact": 16, "mst_compact_log2": 22,
            "mst_split_ratio": 1.5, "cc_block": 256, "cc_ept": 16, "cc_ept_filter": 8,
            "cc_block_filter": 1024}
        P.set_tuning(mst_light_factor=2.0, mst_block=256, cc_ept=32,
                     cpu_share=0.25)  # the others unchanged
        t = P.get_tuning()
        assert t["blocks_per_cu"] == 8 and t["mst_light_factor"] == 2.0
        assert (t["mst_block"], t["mst_ept"], t["cc_ept"]) == (256, 4, 32)
        assert t["cpu_share"] == 0.25
        assert old["mst_light_factor_skewed"] == 0.625  # built-in
        P.set_tuning(mst_light_factor_skewed=0.0)
        assert P.get_tuning()["mst_light_factor_skewed"] == 0.0
        t = P.get_tuning()
        for bad in ({"blocks_per_cu": 65}, {"mst_block": 384}, {"mst_ept": 32},
                    {"mst_ept_compact": 4}, {"mst_compact_log2": 41},
                    {"mst_block": 512, "mst_ept": 16},  # 8192 edges per tile: beyond 4096
                    {"mst_split_ratio": 1.0}, {"cc_block": 1024}, {"cc_ept": 4},
                    {"cc_ept_filter": 32}, {"cc_block_filter": 2048},
                    {"cc_block_filter": 1024, "cc_ept_filter": 16}, {"cpu_share": -0.5},
                    {"mst_light_factor_skewed": -1.0}, {"no_such_knob": 1}):
            with pytest.raises(P.PhemError):
                P.set_tuning(**bad)
        assert P.get_tuning() == t  # refused settings change nothing
    finally:
        P.set_tuning(**old)


def test_nearest_and_cache(tmp_path):
    path = str(tmp_path / "tune.json")
    mk = lambda dev, n, m, bpc: {"device": dev, "size_class": T.size_class(n, m),
                                 "config": {"blocks_per_cu": bpc, "mst_light_factor": 1.0,
                                            "cc_sample_factor": 2.0}}
    T.save(mk("A", 1 << 20, 1 << 24, 8), path)
    T.save(mk("A", 1 << 24, 1 << 28, 16), path)
    T.save(mk("B", 1 << 24, 1 << 28, 4), path)
    T.save(mk("A", 1 << 24, 1 << 28, 32), path)  # replaces the same class
    assert len(T._read(path)) == 3
    assert T.load(1 << 23, 1 << 27, path, apply=False, device="A")["blocks_per_cu"] == 32
    assert T.load(1 << 19, 1 << 23, path, apply=False, device="A")["blocks_per_cu"] == 8
    assert T.load(1 << 24, 1 << 28, path, apply=False, device="B")["blocks_per_cu"] == 4
    assert T.load(1 << 24, 1 << 28, path, apply=False, device="C") is None
    assert T.size_class(1 << 10, 3 << 10) == [10.0, round(math.log2(3 << 10), 1)]
