"""Property tests (hypothesis) of the library's host backend through the C
ABI (phem_cpu_msf, phem_cpu_cc) on small random multigraphs with heavy
weight ties, self-loops and out-of-range endpoints.  The invariants are
checked from first principles, not against the backend's own algorithm:

- |F| = n - c, where c counts the components of the valid edges;
- F is acyclic and every forest edge is a valid input edge;
- cycle property under the (weight, id) key: every valid edge left out of F
  joins two vertices of one tree of F and is heavier than every edge on the
  tree path between them;
- labels are the minimum vertex id of each component;
- the forest is the oracle's Kruskal forest and the canonical Boruvka
  restatement's forest.
"""
import numpy as np
from hypothesis import given, settings, strategies as st

import oracle as O
import phem_native as P

F32_TIES = np.array([0.0, 0.25, 0.5, 0.75, 1.0], np.float32).view(np.uint32)


@st.composite
def graphs(draw):
    n = draw(st.integers(1, 24))
    m = draw(st.integers(0, 60))
    wmode = draw(st.sampled_from([0, 1]))
    # endpoints: mostly in range, a few out of range (ignored by the path)
    ends = st.one_of(st.integers(0, n - 1), st.integers(0, n - 1),
                     st.integers(0, n - 1), st.integers(n, n + 3))
    u = np.array(draw(st.lists(ends, min_size=m, max_size=m)), np.uint32)
    v = np.array(draw(st.lists(ends, min_size=m, max_size=m)), np.uint32)
    if wmode == 1:
        w = F32_TIES[np.array(draw(st.lists(st.integers(0, 4), min_size=m, max_size=m)),
                              np.int64)] if m else np.zeros(0, np.uint32)
    else:
        w = np.array(draw(st.lists(st.integers(0, 3), min_size=m, max_size=m)), np.uint32)
    return n, u, v, np.asarray(w, np.uint32), wmode


def keys(w, wmode):
    return (O.wkey(w, wmode).astype(np.uint64) << np.uint64(32)) | \
        np.arange(len(w), dtype=np.uint64)


def components(n, u, v):
    """Min-id labels by repeated relaxation (independent of any union-find)."""
    lab = np.arange(n, dtype=np.int64)
    ok = (u < n) & (v < n)
    a, b = u[ok].astype(np.int64), v[ok].astype(np.int64)
    changed = True
    while changed:
        changed = False
        for x, y in zip(a, b):
            lo = min(lab[x], lab[y])
            if lab[x] != lo or lab[y] != lo:
                lab[x] = lab[y] = lo
                changed = True
    return lab.astype(np.uint32)


def tree_path_max(n, fu, fv, fk, s, t):
    """Largest key on the forest path s..t (None if s, t are in different
    trees), by depth-first search over the tiny forest."""
    adj = [[] for _ in range(n)]
    for a, b, k in zip(fu, fv, fk):
        adj[a].append((b, k))
        adj[b].append((a, k))
    stack = [(s, -1, 0)]
    seen = {s}
    while stack:
        x, _, best = stack.pop()
        if x == t:
            return best
        for y, k in adj[x]:
            if y not in seen:
                seen.add(y)
                stack.append((y, x, max(best, int(k))))
    return None


@settings(max_examples=300, deadline=None)
@given(graphs())
def test_host_forest_invariants(g):
    n, u, v, w, wmode = g
    ids = P.cpu_msf(n, u, v, w, wmode, 1)
    lab = P.cpu_cc(n, u, v, 1)
    want_lab = components(n, u, v)
    assert np.array_equal(lab, want_lab)
    valid = (u < n) & (v < n) & (u != v)
    c = len(np.unique(want_lab))
    assert len(ids) == n - c
    assert np.all(np.diff(ids.astype(np.int64)) > 0)
    assert valid[ids].all()
    # acyclic: a union-find over the forest edges never closes a cycle
    par = list(range(n))

    def find(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x

    for e in ids:
        ra, rb = find(int(u[e])), find(int(v[e]))
        assert ra != rb
        par[max(ra, rb)] = min(ra, rb)
    # cycle property for every valid non-forest edge
    k = keys(w, wmode)
    infor = np.zeros(len(u), bool)
    infor[ids] = True
    fu, fv, fk = u[ids].astype(np.int64), v[ids].astype(np.int64), k[ids]
    for e in np.nonzero(valid & ~infor)[0]:
        pm = tree_path_max(n, fu, fv, fk, int(u[e]), int(v[e]))
        assert pm is not None and int(k[e]) > pm
    # the same forest as the oracle's Kruskal and canonical Boruvka
    assert np.array_equal(ids, O.kruskal(n, u, v, w, wmode)[0])
    assert np.array_equal(ids, O.boruvka(n, u, v, w, wmode)["ids"])
    # thread count does not change the answer
    assert np.array_equal(P.cpu_msf(n, u, v, w, wmode, 4), ids)
