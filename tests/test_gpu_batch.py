"""GPU parity of the batched solve (phem_batch_solve): k independent graphs
solved as one disjoint union give, for every graph, exactly the forest ids,
min-id labels and exact weight the oracle's Kruskal gives for that graph
alone, including graphs with out-of-range local ids, empty graphs and graphs
with no edges."""
import numpy as np
import pytest

import cases
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


def concat(graphs):
    voff, eoff = [0], [0]
    U, V, W = [], [], []
    for n, u, v, w in graphs:
        voff.append(voff[-1] + n)
        eoff.append(eoff[-1] + len(u))
        U.append(u), V.append(v), W.append(w)
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0, np.uint32)
    return voff, eoff, cat(U), cat(V), cat(W)


def check_batch(P, graphs, wmode, light_factor=None):
    voff, eoff, u, v, w = concat(graphs)
    print(f"batch k={len(graphs)} n={voff[-1]} m={eoff[-1]} wmode={wmode} "
          f"lf={light_factor}", flush=True)
    out = P.batch_solve(voff, eoff, dev(u), dev(v), dev(w), wmode,
                        light_factor=light_factor)
    for g, (n, gu, gv, gw) in enumerate(graphs):
        k_ids, k_lab = O.kruskal(n, gu, gv, gw, wmode)
        ids = out["ids"][g].numpy().view(np.uint32)
        assert np.array_equal(ids, k_ids), f"graph {g}"
        assert np.array_equal(out["labels"][g].numpy().view(np.uint32), k_lab), f"graph {g}"
        ws, exact = O.weight_sum(k_ids, gw, wmode)
        assert (out["inexact"][g] == 0) == exact, f"graph {g}"
        if exact:
            assert out["weight_fx"][g] == ws, f"graph {g}"
    assert out["stats"]["mst_count"] == sum(out["counts"])
    return out


def graphs_of(names):
    gs, modes = [], set()
    for name in names:
        n, u, v, w, wmode = cases.make(name)
        gs.append((n, u, v, w))
        modes.add(wmode)
    assert len(modes) == 1
    return gs, modes.pop()


FLOAT_CASES = [c for c in cases.CASES if cases.make(c)[4] == 1]


@pytest.mark.parametrize("lf", [None, 0.0])
def test_batch_of_all_float_cases(P, lf):
    gs, wmode = graphs_of(FLOAT_CASES)
    check_batch(P, gs, wmode, lf)


def test_batch_u32_with_empty_graphs(P):
    n, u, v, w, wmode = cases.make("u32_weights")
    z = np.zeros(0, np.uint32)
    gs = [(0, z, z, z), (n, u, v, w), (7, z, z, z), (0, z, z, z), (n, u[::-1].copy(),
          v[::-1].copy(), w[::-1].copy()), (0, z, z, z)]
    check_batch(P, gs, wmode)


def test_batch_local_ids_do_not_leak(P):
    """Endpoints at or beyond a graph's own vertex count are invalid edges of
    that graph, even though they are valid global ids of the next graph."""
    n = 50
    u = np.array([0, 1, 49, 50, 60, 3], np.uint32)
    v = np.array([1, 50, 0, 2, 61, 3], np.uint32)
    w = cases.f32([0.5, 0.1, 0.2, 0.05, 0.01, 0.0])
    gs = [(n, u, v, w), (n, u.copy(), v.copy(), w.copy()), (100, u.copy(), v.copy(), w.copy())]
    out = check_batch(P, gs, 1)
    assert out["counts"][:2] == [2, 2]
    assert out["counts"][2] == 5


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_random_mix(P, seed):
    """A served-style mix: many small graphs of different shapes and sizes."""
    rng = np.random.default_rng(seed)
    gs = []
    for g in range(64):
        kind = g % 4
        if kind == 0:
            sc = int(rng.integers(4, 12))
            u, v, w = S.rmat(sc, int(rng.integers(1, 8)) << sc, S.stream_seed(seed, g),
                             positive=True)
            gs.append((1 << sc, u, v, w))
        elif kind == 1:
            n = int(rng.integers(1, 3000))
            u, v, w = S.uniform(n, int(rng.integers(0, 4 * n)), S.stream_seed(seed, g))
            gs.append((n, u, v, w))
        elif kind == 2:
            side = int(rng.integers(2, 40))
            u, v, w = S.grid(side, S.stream_seed(seed, g))
            gs.append((side * side, u, v, w))
        else:
            n = int(rng.integers(2, 4000))
            u, v, w = S.path_chords(n, int(rng.integers(0, n)), S.stream_seed(seed, g))
            gs.append((n, u, v, w))
    check_batch(P, gs, 1)


@pytest.mark.parametrize("lf", [None, 0.25, 4.0])
def test_batch_split_per_graph(P, lf):
    """Graphs of very different density in one batch: each gets its own light
    threshold and its own giant component in the heavy-edge filter.  Sparse
    uniform graphs (average degree 2..4), dense ones (degree 32), an RMAT
    graph, a grid with many components, empty graphs in between."""
    z = np.zeros(0, np.uint32)
    gs = []
    for g, (n, d) in enumerate([(1 << 14, 2), (1 << 13, 32), (5000, 4), (1 << 12, 16)]):
        u, v, w = S.uniform(n, n * d // 2, S.stream_seed(7, g), positive=True)
        gs.append((n, u, v, w))
        gs.append((3, z, z, z))
    u, v, w = S.rmat(13, 12 << 13, S.stream_seed(7, 9), positive=True)
    gs.append((1 << 13, u, v, w))
    u, v, w = S.grid(90, S.stream_seed(7, 10))
    gs.append((90 * 90, u, v, w))
    out = check_batch(P, gs, 1, lf)
    st = out["stats"]
    assert st["light_edges"] + st["heavy_edges"] > 0


@pytest.mark.parametrize("lf", [None, 0.0])
def test_batch_maximum_forest(P, lf):
    """PHEM_MST_MAXIMUM through the batched solve: every graph's forest is its
    own maximum spanning forest ((weight, id) descending)."""
    from max_forest import max_forest_oracle
    gs = []
    for g, (n, d) in enumerate([(3000, 3), (1 << 12, 16), (700, 40)]):
        u, v, w = S.uniform(n, n * d // 2, S.stream_seed(11, g), positive=True)
        gs.append((n, u, v, w))
    u, v, w = S.rmat(11, 8 << 11, S.stream_seed(11, 5), positive=True)
    gs.append((1 << 11, u, v, w))
    voff, eoff, cu, cv, cw = concat(gs)
    out = P.batch_solve(voff, eoff, dev(cu), dev(cv), dev(cw), 1,
                        light_factor=lf, maximum=True)
    for g, (n, gu, gv, gw) in enumerate(gs):
        k_ids, k_lab = max_forest_oracle(n, gu, gv, gw, 1)
        assert np.array_equal(out["ids"][g].numpy().view(np.uint32), k_ids), f"graph {g}"
        assert np.array_equal(out["labels"][g].numpy().view(np.uint32), k_lab), f"graph {g}"
        ws, exact = O.weight_sum(k_ids, gw, 1)
        assert exact and out["weight_fx"][g] == ws, f"graph {g}"


def test_batch_single_graph_equals_solo(P):
    n, u, v, w, wmode = cases.make("rmat_small")
    out = check_batch(P, [(n, u, v, w)], wmode)
    ids, lab, st = P.mst_solve(n, dev(u), dev(v), dev(w), wmode)
    assert np.array_equal(out["ids"][0].numpy(), ids.cpu().numpy())
    assert out["weight_fx"][0] == st["weight_fx"]


def test_batch_bad_offsets(P):
    u = dev(np.zeros(4, np.uint32))
    with pytest.raises(P.PhemError):
        P.batch_solve([0, 5, 3], [0, 2, 4], u, u, u)
    with pytest.raises(P.PhemError):
        P.batch_solve([1, 5], [0, 4], u, u, u)
