"""GPU parity of the recorded batched solve (phem_batch_graph_*): one HIP
graph recorded for capacities is replayed for batches of different counts
and sizes (padded to the capacities), with eager library work between the
replays; every replay gives, for every graph, the oracle's forest ids,
min-id labels and exact weight.  Too few recorded rounds are reported as
PHEM_ENOCONV; a batch beyond the capacities is refused."""
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


def make_batch(seed, k, nmax):
    rng = np.random.default_rng(seed)
    out = []
    for g in range(k):
        kind = g % 3
        if kind == 0:
            sc = int(rng.integers(2, max(3, int(np.log2(nmax)))))
            u, v, w = S.rmat(sc, int(rng.integers(0, 12)) << sc, S.stream_seed(seed, g))
            n = 1 << sc
        elif kind == 1:
            n = int(rng.integers(1, nmax))
            u, v, w = S.uniform(n, int(rng.integers(0, 10 * n)), S.stream_seed(seed, g))
        else:
            side = int(rng.integers(2, 40))
            u, v, w = S.grid(side, S.stream_seed(seed, g))
            n = side * side
        out.append((n, u, v, w))
    return out


def check(res, graphs):
    for g, (n, u, v, w) in enumerate(graphs):
        k_ids, k_lab = O.kruskal(n, u, v, w, 1)
        assert np.array_equal(res["ids"][g].numpy().view(np.uint32), k_ids), g
        assert np.array_equal(res["labels"][g].numpy().view(np.uint32), k_lab), g
        ws, exact = O.weight_sum(k_ids, w, 1)
        assert res["weight_fx"][g] == ws and (res["inexact"][g] == 0) == exact, g


@pytest.mark.parametrize("light_factor", [None, 0.0])
def test_replays_of_different_batches(P, light_factor):
    bg = P.BatchGraph(64, 400_000, 3_000_000, P.W_F32, rounds=14, light_factor=light_factor)
    batches = [make_batch(1, 40, 20000), make_batch(2, 7, 60000), make_batch(3, 1, 200000),
               make_batch(4, 63, 3000), []]
    batches[-1] = batches[0][:5]
    for i, graphs in enumerate(batches):
        n_tot = sum(g[0] for g in graphs)
        m_tot = sum(len(g[1]) for g in graphs)
        print(f"replay {i}: {len(graphs)} graphs, {n_tot} vertices, {m_tot} edges", flush=True)
        assert n_tot <= 400_000 and m_tot <= 3_000_000
        dg = [(n, dev(u), dev(v), dev(w)) for n, u, v, w in graphs]
        bg.launch(dg)
        res = bg.results()
        check(res, graphs)
        # eager library work between replays
        n, u, v, w = graphs[0]
        ids, _, _ = P.mst_solve(n, dev(u), dev(v), dev(w), P.W_F32)
        torch.cuda.synchronize()
        assert np.array_equal(ids.cpu().numpy().view(np.uint32), O.kruskal(n, u, v, w, 1)[0])
    bg.close()


def test_too_few_rounds_and_capacity(P):
    # a long path needs many rounds: one recorded round per phase cannot do
    n = 5000
    u = np.arange(n - 1, dtype=np.uint32)
    w = S.weights(3, np.arange(n - 1, dtype=np.uint64))
    graphs = [(n, u, u + 1, w)]
    bg = P.BatchGraph(4, 8000, 20000, P.W_F32, rounds=1)
    bg.launch([(n, dev(u), dev(u + 1), dev(w))])
    with pytest.raises(P.PhemError) as e:
        bg.results()
    assert e.value.code == P.ENOCONV
    # beyond the capacities: refused before any work
    big = [(n, dev(u), dev(u + 1), dev(w))] * 2
    with pytest.raises(P.PhemError) as e:
        bg.launch(big)  # 10000 vertices > 8000
    assert e.value.code == P.EINVAL
    four = [(10, dev(u[:9]), dev(u[:9] + 1), dev(w[:9]))] * 4
    with pytest.raises(P.PhemError) as e:
        bg.launch(four)  # 4 graphs: no slot left for the padding
    assert e.value.code == P.EINVAL
    # enough rounds: the same recorded capacities solve it
    bg2 = P.BatchGraph(4, 8000, 20000, P.W_F32, rounds=20)
    bg2.launch([(n, dev(u), dev(u + 1), dev(w))])
    check(bg2.results(), graphs)
    bg.close()
    bg2.close()
