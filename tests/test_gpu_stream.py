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


NO_EDGE = np.uint64(0xFFFFFFFFFFFFFFFF)


def check_cc_forest(st, n, au, av, labels):
    """The stream's union-find forest: n - c packed records, each one an
    inserted edge, with the partition of every edge so far (so acyclic)."""
    fo = st.cc_forest().cpu().numpy().view(np.uint64)
    f = fo[fo != NO_EDGE]
    assert len(f) == n - len(np.unique(labels))
    lo, hi = np.minimum(au, av).astype(np.uint64), np.maximum(au, av).astype(np.uint64)
    rec = np.unique((lo << np.uint64(32)) | hi)
    pos = np.minimum(np.searchsorted(rec, f), len(rec) - 1)
    assert (rec[pos] == f).all()
    fu = (f >> np.uint64(32)).astype(np.uint32)
    fv = (f & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(O.bfs_labels(n, fu, fv), labels)


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
        check_cc_forest(st, n, au, av, k_lab)
    st.close()


def test_stream_edge_cases(P):
    """Batches that are empty, self-loops only, out of range, re-inserted
    copies of forest edges (lighter, equal and heavier) and equal-weight
    ties: after every batch the forest is Kruskal's over every edge inserted
    so far (invalid records count as inserted, keep their ids and are
    ignored)."""
    n = 3000
    rng = np.random.default_rng(3)
    z = np.zeros(0, np.uint32)
    u0, v0, w0 = S.uniform(n, 4 * n, S.stream_seed(9, 0), positive=True)
    batches = [(u0, v0, w0), (z, z, z)]
    loops = rng.integers(0, n, 500).astype(np.uint32)
    batches.append((loops, loops.copy(), w0[:500].copy()))
    bad_u = rng.integers(0, 2 * n, 700).astype(np.uint32)
    bad_v = (bad_u + n).astype(np.uint32)
    batches.append((bad_u, bad_v, w0[:700].copy()))
    st = P.Stream(n, 4 * n, wmode=1)
    U, V, W = [], [], []

    def check(t):
        au, av, aw = np.concatenate(U), np.concatenate(V), np.concatenate(W)
        k_ids, k_lab = O.kruskal(n, au, av, aw, 1)
        fu, fv, fw, fg = st.forest()
        torch.cuda.synchronize()
        assert np.array_equal(fg.cpu().numpy().astype(np.uint64),
                              k_ids.astype(np.uint64)), f"batch {t}"
        assert np.array_equal(host32(st.labels()), k_lab), f"batch {t}"
        check_cc_forest(st, n, au, av, k_lab)
        s = st.stats()
        ws, exact = O.weight_sum(k_ids, aw, 1)
        # halved weights leave the 2^-24 grid: then no exact sum exists
        assert bool(s["weight_exact"]) == exact, f"batch {t}"
        if exact:
            assert s["weight_fx"] == ws, f"batch {t}"
        return fu, fv, fw

    t = 0
    for u, v, w in batches:
        st.apply(dev(u), dev(v), dev(w))
        U.append(u), V.append(v), W.append(w)
        check(t)
        t += 1
    # copies of current forest edges: lighter, equal and heavier weights
    fu, fv, fw = (host32(x) for x in check(t - 1))
    k = min(len(fu), 600)
    fwf = fw[:k].view(np.float32)
    for scale in (0.5, 1.0, 2.0):
        w = (fwf * np.float32(scale)).astype(np.float32).view(np.uint32)
        u, v = fv[:k].copy(), fu[:k].copy()  # reversed endpoints
        st.apply(dev(u), dev(v), dev(w))
        U.append(u), V.append(v), W.append(w)
        check(t)
        t += 1
    # equal weights everywhere: ties broken by insertion order
    u, v, _ = S.uniform(n, 2000, S.stream_seed(9, 1))
    w = np.full(2000, np.float32(0.25)).view(np.uint32)
    st.apply(dev(u), dev(v), dev(w))
    U.append(u), V.append(v), W.append(w)
    check(t)
    st.close()


def stream_state(st):
    fu, fv, fw, fg = st.forest()
    torch.cuda.synchronize()
    s = st.stats()
    return (host32(fu).copy(), host32(fv).copy(), host32(fw).copy(),
            fg.cpu().numpy().copy(), host32(st.labels()).copy(),
            st.cc_forest().cpu().numpy().copy(),
            {k: s[k] for k in ("mst_count", "weight_fx", "edges_inserted", "batches")})


def same_state(a, b, cc_forest=True):
    """Equal forests, labels and statistics; the union-find forests too when
    cc_forest (the edge that hooks a root depends on thread timing, so after
    further batches only their invariants are comparable)."""
    k = 6 if cc_forest else 5
    return all(np.array_equal(x, y) for x, y in zip(a[:k], b[:k])) and a[6] == b[6]


def test_stream_snapshot_resume(P):
    """Checkpoint and resume: a stream restored from a host snapshot (with
    the original destroyed) and the original continued in parallel reach the
    same forest, labels, union-find forest and statistics after further
    batches, and both equal the oracle."""
    n = 1 << 12
    batches = [S.rmat(12, 1 << 14, 4, positive=True)]
    for b in range(8):
        batches.append(S.uniform(n, 1 << 10, S.stream_seed(6, b)))
    st = P.Stream(n, 1 << 14)
    for u, v, w in batches[:4]:
        st.apply(dev(u), dev(v), dev(w))
    snap = st.save()
    st2 = P.Stream.restore(snap, 1 << 14)
    assert same_state(stream_state(st), stream_state(st2))
    for u, v, w in batches[4:6]:
        st.apply(dev(u), dev(v), dev(w))
        st2.apply(dev(u), dev(v), dev(w))
    assert same_state(stream_state(st), stream_state(st2), cc_forest=False)
    snap2 = st2.save()
    st2.close()
    st3 = P.Stream.restore(snap2, 1 << 10)  # smaller batches from here on
    for u, v, w in batches[6:]:
        st.apply(dev(u), dev(v), dev(w))
        st3.apply(dev(u), dev(v), dev(w))
    a, b = stream_state(st), stream_state(st3)
    assert same_state(a, b, cc_forest=False)
    au = np.concatenate([x[0] for x in batches])
    av = np.concatenate([x[1] for x in batches])
    aw = np.concatenate([x[2] for x in batches])
    k_ids, k_lab = O.kruskal(n, au, av, aw, 1)
    assert np.array_equal(b[3].astype(np.uint64), k_ids.astype(np.uint64))
    assert np.array_equal(b[4], k_lab)
    check_cc_forest(st3, n, au, av, k_lab)
    st.close()
    st3.close()


@pytest.mark.parametrize("n", [1, 5])
def test_stream_snapshot_of_a_fresh_stream(P, n):
    """A snapshot taken before any batch (empty forest) restores to a stream
    that then behaves as a fresh one."""
    st = P.Stream(n, 16)
    st2 = P.Stream.restore(st.save(), 16)
    u = np.array([0, n - 1, 0], np.uint32)
    v = np.array([n - 1, 0, 0], np.uint32)
    w = np.array([0.5, 0.25, 0.75], np.float32).view(np.uint32)
    for s_ in (st, st2):
        s_.apply(dev(u), dev(v), dev(w))
    assert same_state(stream_state(st), stream_state(st2), cc_forest=False)
    k_ids, k_lab = O.kruskal(n, u, v, w, 1)
    assert np.array_equal(stream_state(st2)[3].astype(np.uint64), k_ids.astype(np.uint64))
    st.close()
    st2.close()


def test_stream_restore_rejects_bad_snapshots(P):
    st = P.Stream(1000, 100)
    u, v, w = S.uniform(1000, 100, 3)
    st.apply(dev(u), dev(v), dev(w))
    f = st.stats()["mst_count"]
    snap = st.save()
    st.close()
    bad = snap.copy()
    bad[:8] = 0  # not a snapshot
    truncated = snap[: len(snap) - 1]
    corrupt = snap.copy()
    # the union-find forest record of vertex n - 2 (the second-last 8 bytes):
    # u = 5 > v = 2 is no record the library writes
    corrupt[-16:-8] = np.array([(5 << 32) | 2], np.uint64).view(np.uint8)
    # forest global ids out of order (the first two swapped): the tie order
    # of later batches depends on ascending ids
    assert f >= 2
    head = len(snap) - 20 * f - 12 * 1000
    g0 = head + 12 * f
    unordered = snap.copy()
    unordered[g0:g0 + 8], unordered[g0 + 8:g0 + 16] = snap[g0 + 8:g0 + 16], snap[g0:g0 + 8]
    assert not np.array_equal(unordered, snap)
    for b in (bad, truncated, corrupt, unordered, snap[:10]):
        with pytest.raises(P.PhemError) as e:
            P.Stream.restore(b, 100)
        assert e.value.code == P.EINVAL


def test_stream_rejects_oversized_batch(P):
    st = P.Stream(100, 10)
    u = dev(np.zeros(11, np.uint32))
    with pytest.raises(P.PhemError):
        st.apply(u, u, u)
    st.close()


@pytest.mark.slow
def test_stream_rmat_2e22_256_batches_vs_recompute(P):
    """RMAT n=2^22, m=2^24 (weights [0,1)), then 256 batches of 2^20 uniform
    insertions: at checkpoints the resident forest equals the OpenMP PBBS
    restatement's forest of every edge so far (ids = insertion order), with
    the same exact weight, and the labels equal the CPU union-find's."""
    scale, m0, nb, mb = 22, 1 << 24, 256, 1 << 20
    n = 1 << scale
    print("generating rmat 2^22 / 2^24 and 256 batches", flush=True)
    base = P.gen_rmat(scale, m0, 4)
    st = P.Stream(n, m0, wmode=1)
    st.apply(*base)
    U, V, W = [host32(t).copy() for t in base]
    U, V, W = [U], [V], [W]
    checks = {1, 128, nb}
    for b in range(1, nb + 1):
        bu, bv, bw = P.gen_uniform(n, mb, S.stream_seed(4, b - 1))
        st.apply(bu, bv, bw)
        U.append(host32(bu).copy()), V.append(host32(bv).copy()), W.append(host32(bw).copy())
        if b in checks:
            au, av, aw = np.concatenate(U), np.concatenate(V), np.concatenate(W)
            print(f"batch {b}: checking {len(au)} edges", flush=True)
            p_ids = O.par_mst(n, au, av, aw, 1)
            fu, fv, fw, fg = st.forest()
            gid = fg.cpu().numpy().astype(np.uint64)
            assert np.array_equal(gid, p_ids.astype(np.uint64)), f"batch {b}"
            s = st.stats()
            ws, exact = O.weight_sum(p_ids, aw, 1)
            assert exact and s["weight_fx"] == ws
            assert np.array_equal(host32(st.labels()), O.par_cc(n, au, av)), f"batch {b}"
    st.close()
