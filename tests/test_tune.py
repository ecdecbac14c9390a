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
        assert P.get_tuning() == {"blocks_per_cu": 8, "mst_light_factor": 0.5,
                                  "cc_sample_factor": 3.0}
        P.set_tuning(mst_light_factor=2.0)  # the others unchanged
        assert P.get_tuning()["blocks_per_cu"] == 8
        assert P.get_tuning()["mst_light_factor"] == 2.0
        with pytest.raises(P.PhemError):
            P.set_tuning(blocks_per_cu=65)
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
