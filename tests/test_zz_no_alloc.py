"""phem_mst_solve_noalloc (static scratch, no allocation) against
phem_mst_solve (caller's workspace) through the C ABI: two calls in a row on
the same device, the second on a smaller graph, so the second runs on
scratch that still holds the first call's state; forest ids, labels and
every statistic must be equal, and equal to the oracle's forest."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    import phem_native
    phem_native.require_device()
    return phem_native


def _solve_both(P, n, u, v, w, light_factor):
    ids_a, lab_a, st_a = P.mst_solve(n, u, v, w, P.W_F32, light_factor=light_factor)
    ids_b = torch.full((n,), -1, dtype=torch.int32, device=u.device)
    lab_b = torch.full((n,), -1, dtype=torch.int32, device=u.device)
    st_b = P.mst_solve_noalloc(n, u, v, w, ids_b, lab_b, P.W_F32, light_factor=light_factor)
    torch.cuda.synchronize()
    return (ids_a.cpu().numpy(), lab_a.cpu().numpy(), st_a,
            ids_b[: st_b["mst_count"]].cpu().numpy(), lab_b.cpu().numpy(), st_b)


@pytest.mark.parametrize("light_factor", [None, 0.0])
def test_noalloc_equals_workspace_solve_twice(P, light_factor):
    graphs = [("rmat", 16, 1 << 20, 2), ("uniform", 12, 1 << 15, 5)]
    for kind, scale, m, seed in graphs:
        n = 1 << scale
        print(f"noalloc {kind} n={n} m={m} lf={light_factor}", flush=True)
        if kind == "rmat":
            u, v, w = P.gen_rmat(scale, m, seed, positive=True, device="cuda:0")
        else:
            u, v, w = P.gen_uniform(n, m, seed, device="cuda:0")
        ia, la, sa, ib, lb, sb = _solve_both(P, n, u, v, w, light_factor)
        assert np.array_equal(ia, ib)
        assert np.array_equal(la, lb)
        assert sa == sb
        hu, hv, hw = (t.cpu().numpy().view(np.uint32) for t in (u, v, w))
        k_ids, k_lab = O.kruskal(n, hu, hv, hw, 1)
        assert np.array_equal(ib.view(np.uint32), k_ids)
        assert np.array_equal(lb.view(np.uint32), k_lab)
        assert sb["weight_fx"] == O.weight_sum(k_ids, hw, 1)[0]


def test_noalloc_refuses_oversized(P):
    u = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    ids = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    with pytest.raises(P.PhemError) as e:
        P.lib()  # bound
        st = P.MstStats()
        import ctypes
        P._check(P.lib().phem_mst_solve_noalloc(1 << 17, 1 << 21, P._p(u), P._p(u), P._p(u),
                                                1, P._p(ids), None, None,
                                                ctypes.byref(st), P._stream()),
                 "phem_mst_solve_noalloc")
    assert e.value.code == P.ENOMEM
