"""GPU parity of the Boruvka MST path (phem_mst_solve through the C ABI)
against the CPU oracle: bit-exact forest id sets, min-id labels, exact
weights and the per-round (m_r, c_r) log."""
import numpy as np
import pytest

import cases
import max_forest as MF
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


def host(t):
    return t.cpu().numpy().view(np.uint32)


def run_mst(P, n, u, v, w, wmode, light_factor=None):
    ids, lab, st = P.mst_solve(n, dev(u), dev(v), dev(w), wmode, light_factor=light_factor)
    torch.cuda.synchronize()
    return host(ids), host(lab), st


@pytest.mark.parametrize("name", list(cases.CASES))
def test_mst_canonical_rounds_match_oracle(P, name):
    """Canonical Boruvka (no light/heavy split): forest, labels, weight and
    the per-round (alive edges, active components) log equal the oracle's
    restatement of the reference round loop."""
    n, u, v, w, wmode = cases.make(name)
    print(f"case {name}: n={n} m={len(u)}", flush=True)
    ids, lab, st = run_mst(P, n, u, v, w, wmode, light_factor=0.0)
    k_ids, k_lab = O.kruskal(n, u, v, w, wmode)
    assert np.array_equal(ids, k_ids)
    assert np.array_equal(lab, k_lab)
    b = O.boruvka(n, u, v, w, wmode)
    assert st["rounds"] == b["rounds"]
    assert st["m_log"] == b["log_m"].tolist()
    assert st["c_log"] == b["log_c"].tolist()
    ws, exact = O.weight_sum(k_ids, w, wmode)
    assert st["weight_exact"] == exact
    if exact:
        assert st["weight_fx"] == ws


@pytest.mark.parametrize("lf", [None, 0.25, 1.0, 4.0])
@pytest.mark.parametrize("name", list(cases.CASES))
def test_mst_split_matches_oracle(P, name, lf):
    """Light/heavy split (default path): same forest, labels and weight."""
    n, u, v, w, wmode = cases.make(name)
    print(f"case {name}: n={n} m={len(u)} light_factor={lf}", flush=True)
    ids, lab, st = run_mst(P, n, u, v, w, wmode, light_factor=lf)
    k_ids, k_lab = O.kruskal(n, u, v, w, wmode)
    assert np.array_equal(ids, k_ids)
    assert np.array_equal(lab, k_lab)
    ws, exact = O.weight_sum(k_ids, w, wmode)
    assert st["weight_exact"] == exact
    if exact:
        assert st["weight_fx"] == ws


def test_light_factor_follows_degree_skew(P):
    """A solve that leaves the light factor to the library takes the skewed
    factor (0.625) for an RMAT graph, whose sampled endpoints repeat (hubs),
    and the default (1.0) for a uniform graph; mst_light_factor_skewed = 0
    turns the choice off.  The light edges the solve reports follow the
    factor (the threshold comes from a 1024-key sample, hence the margins);
    the forest is the oracle's in every case."""
    rm = cases.make("rmat_small")
    un = cases.make("uniform_c1")
    old = P.get_tuning()["mst_light_factor_skewed"]
    try:
        for skew, (n, u, v, w, wmode), lo, hi in (
                (0.625, rm, 0.3, 0.8), (0.625, un, 0.85, 1.2), (0.0, rm, 0.85, 1.2)):
            P.set_tuning(mst_light_factor_skewed=skew)
            ids, lab, st = run_mst(P, n, u, v, w, wmode)
            print(f"skew knob {skew}: n={n} light edges {st['light_edges']}", flush=True)
            assert lo * n <= st["light_edges"] <= hi * n
            k_ids, k_lab = O.kruskal(n, u, v, w, wmode)
            assert np.array_equal(ids, k_ids) and np.array_equal(lab, k_lab)
    finally:
        P.set_tuning(mst_light_factor_skewed=old)


def test_mst_repeat_is_deterministic(P):
    n, u, v, w, wmode = cases.make("rmat_small")
    solver = P.MstSolver(n, len(u))
    du, dv, dw = dev(u), dev(v), dev(w)
    first = None
    for _ in range(3):
        ids, lab, st = solver.solve(du, dv, dw, wmode)
        got = (host(ids).copy(), host(lab).copy(), st["weight_fx"])
        if first is None:
            first = got
        assert np.array_equal(got[0], first[0]) and np.array_equal(got[1], first[1])
        assert got[2] == first[2]


def test_profiler_records_one_kernel_or_all(P):
    """phem_prof_only restricts the event profiler to one kernel; NULL
    restores recording of every kernel."""
    n, u, v, w, wmode = cases.make("rmat_small")
    solver = P.MstSolver(n, len(u))
    du, dv, dw = dev(u), dev(v), dev(w)
    try:
        P.prof_only("k_light")
        P.prof_reset()
        P.prof_enable(True)
        solver.solve(du, dv, dw, wmode)
        P.prof_enable(False)
        one = P.prof_read()
        assert list(one) == ["k_light"] and one["k_light"][0] == 1
        P.prof_only(None)
        P.prof_reset()
        P.prof_enable(True)
        solver.solve(du, dv, dw, wmode)
        P.prof_enable(False)
        every = P.prof_read()
        assert {"k_light", "k_heavy", "k_hook", "k_compact"} <= set(every)
    finally:
        P.prof_enable(False)
        P.prof_only(None)
        P.prof_reset()


def test_device_generators_match_numpy(P):
    n, m = 1 << 14, 1 << 17
    du, dv, dw = P.gen_uniform(n, m, 42)
    u, v, w = S.uniform(n, m, 42)
    assert np.array_equal(host(du), u) and np.array_equal(host(dv), v)
    assert np.array_equal(host(dw), w)
    du, dv, dw = P.gen_uniform(n, 1000, 42, id0=5000, positive=True)
    u, v, w = S.uniform(n, 1000, 42, id0=5000, positive=True)
    assert np.array_equal(host(du), u) and np.array_equal(host(dw), w)
    du, dv, dw = P.gen_rmat(14, m, 7, positive=True)
    u, v, w = S.rmat(14, m, 7, positive=True)
    assert np.array_equal(host(du), u) and np.array_equal(host(dv), v)
    assert np.array_equal(host(dw), w)


def test_mst_rejects_bad_arguments(P):
    u = dev(np.array([0, 1], np.uint32))
    with pytest.raises(P.PhemError):
        P.mst_solve(4, u, u, u, wmode=7)
    solver = P.MstSolver(4, 2)
    big = dev(np.zeros(5, np.uint32))
    with pytest.raises(P.PhemError):
        solver.solve(big, big, big)


@pytest.mark.slow
def test_mst_rmat_full_size_vs_parallel_oracle(P):
    """RMAT n=2^24, m=2^28 (a,b,c,d = .57,.19,.19,.05), weights in (0,1]: the
    device forest equals the OpenMP PBBS restatement's, labels equal the
    device union-find's, |F| = n - c."""
    scale, m = 24, 1 << 28
    n = 1 << scale
    print("generating rmat 2^24 / 2^28", flush=True)
    du, dv, dw = P.gen_rmat(scale, m, 2, positive=True)
    ids, lab, st = P.mst_solve(n, du, dv, dw, P.W_F32)
    torch.cuda.synchronize()
    d_ids = host(ids).copy()
    d_lab = host(lab).copy()
    print(f"device forest {len(d_ids)} edges, rounds {st['rounds']}", flush=True)
    e = P.pack_edges(du, dv)
    clab, _, _ = P.cc_solve(n, e, forest=False)
    assert np.array_equal(host(clab), d_lab)
    comps = int((d_lab == np.arange(n, dtype=np.uint32)).sum())
    assert len(d_ids) == n - comps
    u, v, w = host(du).copy(), host(dv).copy(), host(dw).copy()
    del e, du, dv, dw
    p_ids = O.par_mst(n, u, v, w, 1)
    assert np.array_equal(d_ids, p_ids)
    ws, exact = O.weight_sum(p_ids, w, 1)
    assert exact and st["weight_fx"] == ws


@pytest.mark.slow
def test_mst_cc_rmat_permuted_ids_full_size(P):
    """The benchmark's permuted graph: RMAT n=2^24, m=2^28 with its vertex ids
    permuted on the device (seed 9), so the hubs and the giant component are
    spread over the id range.  The forest, exact weight and labels equal the
    oracle's; the CC solve's labels (its unions then build the giant
    component under roots that change many times) equal them too."""
    scale, m = 24, 1 << 28
    n = 1 << scale
    print("generating rmat 2^24 / 2^28, ids permuted", flush=True)
    du, dv, dw = P.gen_rmat(scale, m, 2, positive=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    perm = torch.randperm(n, device="cuda", generator=g, dtype=torch.int64).to(torch.int32)
    du, dv = perm[du.long()], perm[dv.long()]
    del perm
    ids, lab, st = P.mst_solve(n, du, dv, dw, P.W_F32)
    e = P.pack_edges(torch.minimum(du, dv), torch.maximum(du, dv))
    clab, _, _ = P.cc_solve(n, e, forest=False)
    torch.cuda.synchronize()
    d_ids, d_lab, c_lab = host(ids).copy(), host(lab).copy(), host(clab).copy()
    u, v, w = host(du).copy(), host(dv).copy(), host(dw).copy()
    del e, du, dv, dw, ids, lab, clab
    print(f"device forest {len(d_ids)} edges, rounds {st['rounds']}", flush=True)
    p_ids = O.par_mst(n, u, v, w, 1)
    assert np.array_equal(d_ids, p_ids)
    ws, exact = O.weight_sum(p_ids, w, 1)
    assert exact and st["weight_fx"] == ws
    want = O.par_cc(n, u, v)
    assert np.array_equal(d_lab, want)
    assert np.array_equal(c_lab, want)


@pytest.mark.parametrize("wkind", ["int16_f32", "int16_u32", "all_equal"])
def test_mst_rmat_tied_weights(P, wkind):
    """RMAT n = 2^18, m = 2^22 whose weights tie: integers 1..16 (as float32
    and as uint32) and one weight for every edge.  Most offers to a hub's
    entry then carry its own weight, so the split schedule's tie-checked
    offers decide by edge id; forest, labels and exact weight equal the
    oracle's in both schedules."""
    n = 1 << 18
    du, dv, dw = P.gen_rmat(18, 1 << 22, 11, positive=True)
    u, v = host(du).copy(), host(dv).copy()
    r = np.random.default_rng(11).integers(1, 17, len(u))
    if wkind == "int16_f32":
        w, wmode = r.astype(np.float32).view(np.uint32), 1
    elif wkind == "int16_u32":
        w, wmode = r.astype(np.uint32), 0
    else:
        w, wmode = np.full(len(u), np.float32(0.5)).view(np.uint32), 1
    k_ids, k_lab = O.kruskal(n, u, v, w, wmode)
    ws, exact = O.weight_sum(k_ids, w, wmode)
    for lf in (None, 0.0):
        ids, lab, st = run_mst(P, n, u, v, w, wmode, light_factor=lf)
        assert np.array_equal(ids, k_ids), lf
        assert np.array_equal(lab, k_lab), lf
        assert exact and st["weight_fx"] == ws, lf
        if lf is None:
            assert st["tau"] != 0xFFFFFFFFFFFFFFFF  # the split schedule ran


@pytest.mark.parametrize("name", ["rmat_small", "grid_small", "out_of_range", "u32_weights"])
def test_mst_aos_layout_matches_soa(P, name):
    n, u, v, w, wmode = cases.make(name)
    aos = np.stack([u, v, w], axis=1).reshape(-1).copy()
    solver = P.MstSolver(n, len(u))
    ids, lab, st = solver.solve_aos(dev(aos), wmode)
    torch.cuda.synchronize()
    k_ids, k_lab = O.kruskal(n, u, v, w, wmode)
    assert np.array_equal(host(ids), k_ids)
    assert np.array_equal(host(lab), k_lab)


def _full_check(P, n, u, v, w, canonical_log=False):
    du, dv, dw = dev(u), dev(v), dev(w)
    ids, lab, st = P.mst_solve(n, du, dv, dw, P.W_F32)
    torch.cuda.synchronize()
    d_ids, d_lab = host(ids).copy(), host(lab).copy()
    p_ids = O.par_mst(n, u, v, w, 1)
    assert np.array_equal(d_ids, p_ids)
    ws, exact = O.weight_sum(p_ids, w, 1)
    assert exact and st["weight_fx"] == ws
    assert np.array_equal(d_lab, O.par_cc(n, u, v))
    clab, _, _ = P.cc_solve(n, P.pack_edges(du, dv), forest=False)
    assert np.array_equal(host(clab), d_lab)
    if canonical_log:
        c_ids, _, cst = P.mst_solve(n, du, dv, dw, P.W_F32, light_factor=0.0)
        assert np.array_equal(host(c_ids), p_ids)
        b = O.boruvka(n, u, v, w, 1)
        assert cst["m_log"] == b["log_m"].tolist()
        assert cst["c_log"] == b["log_c"].tolist()
    return st


@pytest.mark.slow
def test_mst_grid_4096_tie_heavy(P):
    """4096 x 4096 grid, 20% of edges deleted, integer weights 1..16 as
    float32: massive ties resolved by edge id, many components."""
    print("generating grid 4096^2", flush=True)
    u, v, w = S.grid(4096, 3)
    st = _full_check(P, 4096 * 4096, u, v, w)
    print(f"grid: m={len(u)} forest={st['mst_count']} rounds={st['rounds']}", flush=True)


@pytest.mark.slow
def test_mst_path_2e22_chords_deep_rounds(P):
    """Random-permuted path on 2^22 vertices plus 2^20 uniform chords: long
    hook chains and many rounds; the canonical schedule's per-round log
    equals the oracle's restatement of the reference round loop."""
    n = 1 << 22
    print("generating path 2^22 + 2^20 chords", flush=True)
    u, v, w = S.path_chords(n, 1 << 20, 4)
    st = _full_check(P, n, u, v, w, canonical_log=True)
    print(f"path: m={len(u)} forest={st['mst_count']} rounds={st['rounds']}", flush=True)


@pytest.mark.parametrize("lf", [None, 0.0])
@pytest.mark.parametrize("name", list(cases.CASES))
def test_maximum_spanning_forest(P, name, lf):
    n, u, v, w, wmode = cases.make(name)
    ids, lab, st = P.mst_solve(n, dev(u), dev(v), dev(w), wmode, light_factor=lf,
                               maximum=True)
    torch.cuda.synchronize()
    x_ids, x_lab = MF.max_forest_oracle(n, u, v, w, wmode)
    assert np.array_equal(host(ids), x_ids)
    assert np.array_equal(host(lab), x_lab)
    ws, exact = O.weight_sum(x_ids, w, wmode)
    assert st["weight_exact"] == exact
    if exact:
        assert st["weight_fx"] == ws



def giant_above_neighbours(seed):
    """A dense random graph on the upper half of the ids (it becomes the giant
    component of the light phase, so its root G >= n/2) and small components
    on the lower ids, joined inside by light edges and to the upper half only
    by heavy edges: their components x < G reach the giant-adjacent minimum
    path that is kept as an alive edge instead of hooked at once."""
    rng = np.random.default_rng(seed)
    n = 1 << 13
    h = n // 2
    mg = 16 * h
    gu = rng.integers(h, n, mg)
    gv = rng.integers(h, n, mg)
    gw = rng.integers(1, 1 << 20, mg)              # light-ish weights
    cu = rng.integers(0, h, 2 * h)                  # small components below h
    cv = np.minimum(cu + rng.integers(1, 4, 2 * h), h - 1)
    cw = rng.integers(1, 1 << 20, 2 * h)
    au = rng.integers(0, h, 4 * h)                  # attachments: heavy only
    av = rng.integers(h, n, 4 * h)
    aw = rng.integers(1 << 23, 1 << 24, 4 * h)
    u = np.concatenate([gu, cu, au]).astype(np.uint32)
    v = np.concatenate([gv, cv, av]).astype(np.uint32)
    w = (np.concatenate([gw, cw, aw]).astype(np.float32) / np.float32(1 << 24)).view(np.uint32)
    perm = rng.permutation(len(u))  # edge ids in random order
    return n, u[perm].copy(), v[perm].copy(), w[perm].copy()


@pytest.mark.parametrize("lf", [None, 0.5, 2.0])
@pytest.mark.parametrize("seed", [1, 2])
def test_split_giant_above_its_neighbours(P, seed, lf):
    n, u, v, w = giant_above_neighbours(seed)
    ids, lab, st = run_mst(P, n, u, v, w, 1, light_factor=lf)
    k_ids, k_lab = O.kruskal(n, u, v, w, 1)
    assert np.array_equal(ids, k_ids)
    assert np.array_equal(lab, k_lab)
    ws, exact = O.weight_sum(k_ids, w, 1)
    assert exact and st["weight_fx"] == ws
    if lf is None:
        assert st["tau"] != 0xFFFFFFFFFFFFFFFF  # the split schedule ran
        assert st["giant_edges"] > 0           # kept giant edges (x < G)


def test_batch_giant_above_its_neighbours(P):
    graphs = [giant_above_neighbours(s) for s in (3, 4, 5)]
    voff, eoff = [0], [0]
    for n, u, v, w in graphs:
        voff.append(voff[-1] + n)
        eoff.append(eoff[-1] + len(u))
    cat = [torch.from_numpy(np.concatenate([g[i] for g in graphs]).view(np.int32)).cuda()
           for i in (1, 2, 3)]
    r = P.batch_solve(voff, eoff, *cat)
    for g, (n, u, v, w) in enumerate(graphs):
        k_ids, k_lab = O.kruskal(n, u, v, w, 1)
        assert np.array_equal(r["ids"][g].numpy().view(np.uint32), k_ids), g
        assert np.array_equal(r["labels"][g].numpy().view(np.uint32), k_lab), g
