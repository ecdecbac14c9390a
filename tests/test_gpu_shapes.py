"""Every kernel shape the tuning knobs can select (threads per block and
edges / records per thread of the MST input passes, the compaction rounds
and the CC unions, the compaction and split thresholds) gives the oracle's
result: forest ids, labels and exact weight of the MST solve (split and
canonical schedules, one graph and a batch), and the CC labels."""
import itertools

import numpy as np
import pytest

import cases
import oracle as O
import synth as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = ["rmat_small", "grid_small", "path_chords_small", "many_components",
         "duplicates_equal_w", "out_of_range", "uniform_c1"]


@pytest.fixture(scope="module")
def P():
    import phem_native
    phem_native.require_device()
    old = phem_native.get_tuning()
    yield phem_native
    phem_native.set_tuning(**old)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def host(t):
    return t.cpu().numpy().view(np.uint32)


_BIG = {}


def big_graph(P):
    """RMAT n = 2^18, m = 2^22: large enough for the 4096-edge tiles (graphs
    of fewer than 2^22 edges run at most 2048 edges per tile)."""
    if not _BIG:
        u, v, w = P.gen_rmat(18, 1 << 22, 7, positive=True)
        torch.cuda.synchronize()
        _BIG["g"] = (1 << 18, host(u).copy(), host(v).copy(), host(w).copy(), 1)
        _BIG["k"] = O.kruskal(*_BIG["g"])
    return _BIG["g"], _BIG["k"]


def check_mst(P, name, lf):
    if name == "big":
        (n, u, v, w, wmode), (k_ids, k_lab) = big_graph(P)
    else:
        n, u, v, w, wmode = cases.make(name)
        k_ids, k_lab = O.kruskal(n, u, v, w, wmode)
    ids, lab, st = P.mst_solve(n, dev(u), dev(v), dev(w), wmode, light_factor=lf)
    torch.cuda.synchronize()
    assert np.array_equal(host(ids), k_ids), name
    assert np.array_equal(host(lab), k_lab), name
    ws, exact = O.weight_sum(k_ids, w, wmode)
    assert st["weight_exact"] == exact and (not exact or st["weight_fx"] == ws), name


@pytest.mark.parametrize("block,ept", [(256, 4), (256, 8), (256, 16), (512, 4), (512, 8)])
def test_mst_input_pass_shapes(P, block, ept):
    # both compaction tiles: 8 per thread everywhere, 16 from 2^22 edges; the
    # big graph runs the 4096-edge tiles (smaller graphs at most 2048)
    for compact in (8, 16):
        P.set_tuning(mst_block=block, mst_ept=ept, mst_ept_compact=compact)
        print(f"mst shape block={block} ept={ept} compact={compact}", flush=True)
        for name in CASES + ["big"]:
            for lf in (0.25, 1.0, 0.0):
                check_mst(P, name, lf)
    P.set_tuning(mst_block=512, mst_ept=4, mst_ept_compact=16)
    with pytest.raises(P.PhemError):  # 512 x 16 is beyond the largest compiled tile
        P.set_tuning(mst_ept=16)


@pytest.mark.parametrize("ratio", [1.25, 3.0, 40.0])
def test_mst_split_ratio(P, ratio):
    """The split threshold only moves the choice between the two schedules."""
    P.set_tuning(mst_split_ratio=ratio)
    try:
        for name in CASES + ["big"]:
            check_mst(P, name, 1.0)
    finally:
        P.set_tuning(mst_split_ratio=1.5)


@pytest.mark.parametrize("block,ept", [(512, 4), (512, 8), (256, 16)])
def test_batch_with_shapes(P, block, ept):
    """A batch of graphs of different density (per-graph thresholds and
    giants) with a non-default input-pass shape."""
    P.set_tuning(mst_block=block, mst_ept=ept)
    try:
        rng = np.random.default_rng(block + ept)
        graphs = []
        for g in range(12):
            n = int(rng.integers(2, 3000))
            m = int(rng.integers(0, 20 * n))
            graphs.append((n,) + S.uniform(n, m, 100 + g))
        voff = np.concatenate([[0], np.cumsum([g[0] for g in graphs])])
        eoff = np.concatenate([[0], np.cumsum([len(g[1]) for g in graphs])])
        cat = [np.concatenate([g[i] for g in graphs]) for i in (1, 2, 3)]
        r = P.batch_solve(voff.tolist(), eoff.tolist(), *(dev(x) for x in cat), P.W_F32)
        for g, (n, u, v, w) in enumerate(graphs):
            k_ids, k_lab = O.kruskal(n, u, v, w, 1)
            assert np.array_equal(r["ids"][g].numpy().view(np.uint32), k_ids), g
            assert np.array_equal(r["labels"][g].numpy().view(np.uint32), k_lab), g
            assert r["weight_fx"][g] == O.weight_sum(k_ids, w, 1)[0], g
    finally:
        P.set_tuning(mst_block=512, mst_ept=4)


@pytest.mark.parametrize("block,fblock", [(256, 256), (512, 512), (256, 1024)])
def test_cc_union_shapes(P, block, fblock):
    for ept, eptf in itertools.product([8, 16, 32], [4, 8, 16]):
        if fblock == 1024 and eptf > 8:
            continue  # not compiled
        P.set_tuning(cc_block=block, cc_ept=ept, cc_block_filter=fblock,
                     cc_ept_filter=eptf)
        print(f"cc shape block={block} ept={ept} filter block={fblock} ept={eptf}",
              flush=True)
        for name in CASES:
            n, u, v, w, wmode = cases.make(name)
            rec = ((np.minimum(u, v).astype(np.uint64) << np.uint64(32))
                   | np.maximum(u, v).astype(np.uint64))
            lab, _, forest = P.cc_solve(n, torch.from_numpy(rec.view(np.int64)).cuda())
            want = O.bfs_labels(n, u, v)
            assert np.array_equal(host(lab), want), name
            f = forest.cpu().numpy().view(np.uint64)
            assert int((f != np.uint64(P.NO_EDGE)).sum()) == n - len(np.unique(want)), name
    P.set_tuning(cc_block=256, cc_ept=16, cc_block_filter=256, cc_ept_filter=8)
