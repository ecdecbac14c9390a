"""GPU test of the autotuner: it picks values from its grid, every setting it
tries gives the same forest, weight and labels (tune() raises otherwise), and
the tuned solve still equals the oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def P():
    import phem_native
    phem_native.require_device()
    return phem_native


def test_tune_small_graph(P, tmp_path):
    import phem_tune as T
    old = P.get_tuning()
    grid = {"blocks_per_cu": [4, 16], "mst_light_factor": [0.5, 1.0],
            "cc_sample_factor": [1.0, 2.0], "mst_block": [256, 512], "mst_ept": [4, 8],
            "cc_block": [256, 512], "cc_ept": [16], "cc_ept_filter": [8]}
    try:
        r = T.tune(1 << 16, 1 << 20, "rmat", seed=3, reps=2, grid=grid)
        for k, vals in grid.items():
            assert r["config"][k] in vals
        # 2 + 4 (block x ept jointly) + 2 + 2 (cc_block x cc_ept) + 1 (the
        # filtered unions' shape) + 2 settings
        assert len(r["log"]) == 13
        path = str(tmp_path / "tune.json")
        T.save(r, path)
        assert T.load(1 << 16, 1 << 20, path) == r["config"]
        assert P.get_tuning() == r["config"]
        n, m = 1 << 16, 1 << 20
        u, v, w = P.gen_rmat(16, m, 3, positive=True)
        ids, lab, st = P.mst_solve(n, u, v, w, P.W_F32)
        hu, hv, hw = (t.cpu().numpy().view(np.uint32) for t in (u, v, w))
        k_ids, k_lab = O.kruskal(n, hu, hv, hw, 1)
        assert np.array_equal(ids.cpu().numpy().view(np.uint32), k_ids)
        assert np.array_equal(lab.cpu().numpy().view(np.uint32), k_lab)
    finally:
        P.set_tuning(**old)
