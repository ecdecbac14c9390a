"""GPU parity of the union-find connected-components path
(phem_cc_* through the C ABI) against the CPU oracle: min-id labels bit-exact,
spanning-forest invariants (the forest's edge choice depends on the schedule
in the reference too), batched application, and the CPU/GPU forest merge."""
import numpy as np
import pytest

import cases
import oracle as O
import synth as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NO_EDGE = np.uint64(0xFFFFFFFFFFFFFFFF)


@pytest.fixture(scope="module")
def P():
    import phem_native
    phem_native.require_device()
    return phem_native


def packed(u, v):
    """Reference record format: (min << 32) | max (CC/graph.cu:74-77)."""
    a = np.minimum(u, v).astype(np.uint64)
    b = np.maximum(u, v).astype(np.uint64)
    return (a << np.uint64(32)) | b


def forest_ids(n, u, v, forest):
    """Map packed forest records back to edge ids of (u, v) (any matching id)."""
    rec = packed(u, v)
    f = forest[forest != NO_EDGE]
    order = np.argsort(rec, kind="stable")
    pos = np.searchsorted(rec[order], f)
    assert (rec[order][pos] == f).all()
    return np.sort(order[pos]).astype(np.uint32)


def assert_spanning_forest(n, fo, want):
    """fo (n packed slots) holds n - c records whose union-find partition is
    `want`: a graph with n - c edges and c components has no cycle, so the
    records form a spanning forest of every component."""
    f = fo[fo != NO_EDGE]
    c = len(np.unique(want))
    assert len(f) == n - c
    fu = (f >> np.uint64(32)).astype(np.uint32)
    fv = (f & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(O.par_cc(n, fu, fv), want)


@pytest.mark.parametrize("name", list(cases.CASES))
def test_cc_matches_oracle(P, name):
    n, u, v, w, wmode = cases.make(name)
    if n == 0:
        pytest.skip("no vertices")
    print(f"case {name}: n={n} m={len(u)}", flush=True)
    e = torch.from_numpy(packed(u, v).view(np.int64)).cuda()
    lab, par, f = P.cc_solve(n, e)
    torch.cuda.synchronize()
    lab = lab.cpu().numpy().view(np.uint32)
    want = O.bfs_labels(n, u, v)
    assert np.array_equal(lab, want)
    fo = f.cpu().numpy().view(np.uint64)
    ids = forest_ids(n, u, v, fo)  # every forest record is an input record
    assert_spanning_forest(n, fo, want)
    # duplicates may map two slots to one id
    if len(np.unique(ids)) == len(ids):
        assert O.check_forest(n, u, v, ids, want) == 0


def test_cc_batches_equal_one_shot(P):
    n = 1 << 16
    u, v, _ = S.rmat(16, 1 << 20, 3)
    e = torch.from_numpy(packed(u, v).view(np.int64)).cuda()
    par = torch.empty(n, dtype=torch.int32, device="cuda")
    lab = torch.empty(n, dtype=torch.int32, device="cuda")
    P.cc_init(n, par)
    for k in range(0, len(u), 1 << 17):
        P.cc_union(n, e[k:k + (1 << 17)], par)
        P.cc_labels(n, par, lab)
        torch.cuda.synchronize()
        want = O.bfs_labels(n, u[: k + (1 << 17)], v[: k + (1 << 17)])
        assert np.array_equal(lab.cpu().numpy().view(np.uint32), want)


def test_cc_merge_cpu_and_gpu_forests(P):
    """HAST split: the GPU takes the first part of the edges, the CPU the
    rest; the merged forest spans the whole graph (CC/main.cu:174-193)."""
    n = 1 << 15
    u, v, _ = S.uniform(n, 1 << 16, 21)
    split = 40000
    e = torch.from_numpy(packed(u[:split], v[:split]).view(np.int64)).cuda()
    _, _, f1 = P.cc_solve(n, e)
    _, cf = O.par_cc(n, u[split:], v[split:], forest=True)
    extra = np.full(n, NO_EDGE, np.uint64)
    hooked = cf != O.NO_EDGE
    extra[hooked] = packed(u[split:][cf[hooked]], v[split:][cf[hooked]])
    par = torch.empty(n, dtype=torch.int32, device="cuda")
    lab = torch.empty(n, dtype=torch.int32, device="cuda")
    f2 = torch.empty(n, dtype=torch.int64, device="cuda")
    P.cc_merge_forests(n, f1, torch.from_numpy(extra.view(np.int64)).cuda(), par, f2, lab)
    torch.cuda.synchronize()
    want = O.bfs_labels(n, u, v)
    assert np.array_equal(lab.cpu().numpy().view(np.uint32), want)
    fo = f2.cpu().numpy().view(np.uint64)
    assert int((fo != NO_EDGE).sum()) == n - len(np.unique(want))


@pytest.mark.parametrize("fblock", [256, 512, 1024])
def test_cc_filtered_shapes_rmat(P, fblock):
    """Every block size of the filtered unions (LDS summaries of 16, 32 and
    64 KB; at n = 2^23 the summary covers all vertices from 512 threads up)
    and both record counts per thread give the oracle's labels and a spanning
    forest, on a graph large enough for a giant component to form."""
    scale, m = 23, 1 << 26
    n = 1 << scale
    du, dv, _ = P.gen_rmat(scale, m, 5, positive=True)
    e = P.pack_edges(du, dv)
    u, v = du.cpu().numpy().view(np.uint32), dv.cpu().numpy().view(np.uint32)
    want = O.par_cc(n, u, v)
    old = P.get_tuning()
    try:
        for eptf in (4, 8):
            P.set_tuning(cc_block_filter=fblock, cc_ept_filter=eptf)
            print(f"cc filtered block={fblock} ept={eptf}", flush=True)
            lab, _, f = P.cc_solve(n, e)
            torch.cuda.synchronize()
            assert np.array_equal(lab.cpu().numpy().view(np.uint32), want), eptf
            assert_spanning_forest(n, f.cpu().numpy().view(np.uint64), want)
    finally:
        P.set_tuning(**old)


@pytest.mark.parametrize("seed", [9, 10])
def test_cc_rmat_permuted_ids(P, seed):
    """RMAT n = 2^22, m = 2^26 with its vertex ids permuted: the giant
    component's root (its smallest id) changes many times while the sample
    phase builds it, the case the device-scope root check in the unions is
    for.  Labels, spanning forest and the merge of two half forests equal the
    oracle's."""
    scale, m = 22, 1 << 26
    n = 1 << scale
    du, dv, _ = P.gen_rmat(scale, m, 3, positive=True)
    pm = torch.from_numpy(np.random.default_rng(seed).permutation(n).astype(np.int32)).cuda()
    du, dv = pm[du.long()], pm[dv.long()]
    e = P.pack_edges(torch.minimum(du, dv), torch.maximum(du, dv))
    u, v = du.cpu().numpy().view(np.uint32), dv.cpu().numpy().view(np.uint32)
    want = O.par_cc(n, u, v)
    lab, _, f = P.cc_solve(n, e)
    torch.cuda.synchronize()
    assert np.array_equal(lab.cpu().numpy().view(np.uint32), want)
    assert_spanning_forest(n, f.cpu().numpy().view(np.uint64), want)
    # the forests of the two halves of the records, merged
    h = m // 2
    _, _, f1 = P.cc_solve(n, e[:h].contiguous())
    _, _, f2 = P.cc_solve(n, e[h:].contiguous())
    par = torch.empty(n, dtype=torch.int32, device="cuda")
    lab3 = torch.empty(n, dtype=torch.int32, device="cuda")
    f3 = torch.empty(n, dtype=torch.int64, device="cuda")
    P.cc_merge_forests(n, f1, f2, par, f3, lab3)
    torch.cuda.synchronize()
    assert np.array_equal(lab3.cpu().numpy().view(np.uint32), want)
    assert_spanning_forest(n, f3.cpu().numpy().view(np.uint64), want)


@pytest.mark.slow
def test_cc_rmat_full_size(P):
    scale, m = 24, 1 << 28
    n = 1 << scale
    du, dv, _ = P.gen_rmat(scale, m, 2, positive=True)
    e = P.pack_edges(du, dv)
    lab, _, f = P.cc_solve(n, e)
    torch.cuda.synchronize()
    # every forest record is an input record (membership tested on the device)
    fd = f[f != -1]
    assert bool(torch.isin(fd, e).all())
    del fd
    lab = lab.cpu().numpy().view(np.uint32)
    u, v = du.cpu().numpy().view(np.uint32), dv.cpu().numpy().view(np.uint32)
    del e, du, dv
    want = O.par_cc(n, u, v)
    assert np.array_equal(lab, want)
    assert_spanning_forest(n, f.cpu().numpy().view(np.uint64), want)
    # the largest filtered-union shape (64 KB LDS summary over all 2^24 ids)
    old = P.get_tuning()
    try:
        P.set_tuning(cc_block_filter=1024, cc_ept_filter=8)
        e = P.pack_edges(torch.from_numpy(u.view(np.int32)).cuda(),
                         torch.from_numpy(v.view(np.int32)).cuda())
        lab2, _, f2 = P.cc_solve(n, e)
        torch.cuda.synchronize()
        assert np.array_equal(lab2.cpu().numpy().view(np.uint32), want)
        assert_spanning_forest(n, f2.cpu().numpy().view(np.uint64), want)
    finally:
        P.set_tuning(**old)
