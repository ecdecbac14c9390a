"""GPU parity of the streaming forest (phem_stream_*): after every batch the
device forest equals the oracle's Kruskal forest of every edge inserted so
far (ids = insertion order), with the same exact weight, and the union-find
labels equal BFS labels of all edges so far."""
import numpy as np
import pytest

import oracle as O
import synth as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def P():
    import phem_native
    phem_native.require_device()
    return phem_native


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def host32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("wmode", [1, 0])
def test_stream_forest_equals_recompute(P, wmode):
    scale, m0 = 13, 1 << 15
    n = 1 << scale
    batches = [S.rmat(scale, m0, 4, positive=True)]
    for b in range(10):
        batches.append(S.uniform(n, 1 << 11, S.stream_seed(4, b)))
    if wmode == 0:  # integer weights with many ties
        batches = [(u, v, (w.view(np.float32) * 8).astype(np.uint32)) for u, v, w in batches]
    st = P.Stream(n, m0, wmode=wmode)
    U, V, W = [], [], []
    for t, (u, v, w) in enumerate(batches):
        st.apply(dev(u), dev(v), dev(w))
        U.append(u), V.append(v), W.append(w)
        au, av, aw = np.concatenate(U), np.concatenate(V), np.concatenate(W)
        k_ids, k_lab = O.kruskal(n, au, av, aw, wmode)
        fu, fv, fw, fg = st.forest()
        torch.cuda.synchronize()
        gid = fg.cpu().numpy().astype(np.uint64)
        assert np.array_equal(gid, k_ids.astype(np.uint64)), f"batch {t}"
        assert np.array_equal(host32(fu), au[k_ids]) and np.array_equal(host32(fw), aw[k_ids])
        s = st.stats()
        ws, exact = O.weight_sum(k_ids, aw, wmode)
        if exact:
            assert s["weight_fx"] == ws
        assert s["edges_inserted"] == len(au)
        assert np.array_equal(host32(st.labels()), k_lab), f"batch {t}"
    st.close()


def test_stream_rejects_oversized_batch(P):
    st = P.Stream(100, 10)
    u = dev(np.zeros(11, np.uint32))
    with pytest.raises(P.PhemError):
        st.apply(u, u, u)
    st.close()
