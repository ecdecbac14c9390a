"""GPU parity of the recorded solve (phem_mst_graph_*): a whole MST solve
captured into a HIP graph with a fixed number of Boruvka rounds per phase
that end on the device.  Every replay, with eager library work and host
synchronisation between replays, must give the forest, labels, weight and
round log of the eager phem_mst_solve and of the oracle."""
import ctypes

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


def host(t):
    return t.cpu().numpy().view(np.uint32)


LOG_KEYS = ("rounds", "light_rounds", "mst_count", "weight_fx", "weight_exact",
            "valid_edges", "tau", "light_edges", "heavy_edges", "giant_edges",
            "sum_m", "sum_c", "m_log", "c_log")


def check_replay(P, g, n, u, v, w, wmode, eager):
    ids, lab, st = g.result()
    k_ids, k_lab = O.kruskal(n, u, v, w, wmode)
    assert np.array_equal(host(ids), k_ids)
    assert np.array_equal(host(lab), k_lab)
    for k in LOG_KEYS:
        assert st[k] == eager[k], k


@pytest.mark.parametrize("lf", [None, 0.0])
@pytest.mark.parametrize("name", list(cases.CASES))
def test_graph_replays_equal_eager(P, name, lf):
    """rounds=0 (picked by an eager solve): three replays, each after eager
    work on the same stream and a host synchronisation."""
    n, u, v, w, wmode = cases.make(name)
    print(f"case {name}: n={n} m={len(u)} light_factor={lf}", flush=True)
    du, dv, dw = dev(u), dev(v), dev(w)
    _, _, eager = P.mst_solve(n, du, dv, dw, wmode, light_factor=lf)
    g = P.MstGraph(n, du, dv, dw, wmode, rounds=0, light_factor=lf)
    other = P.MstSolver(max(n, 2), max(len(u), 1))
    for rep in range(3):
        g.launch()
        check_replay(P, g, n, u, v, w, wmode, eager)
        # eager work between replays: another solve and a torch kernel
        other.solve(du, dv, dw, wmode, n=n, light_factor=lf)
        (dw.view(torch.float32) * 1.0).sum()
        torch.cuda.synchronize()
    g.close()


@pytest.mark.parametrize("lf", [None, 0.0])
def test_graph_new_weights_between_replays(P, lf):
    """The recorded graph solves what its buffers hold at launch: new weights
    (and endpoints) copied into the same buffers between replays."""
    n, m = 1 << 12, 1 << 15
    u, v, w = S.rmat(12, m, 11, positive=True)
    du, dv, dw = dev(u), dev(v), dev(w)
    g = P.MstGraph(n, du, dv, dw, P.W_F32, rounds=12, light_factor=lf)
    for rep in range(4):
        u2, v2, w2 = S.rmat(12, m, 100 + rep, positive=True)
        du.copy_(dev(u2)), dv.copy_(dev(v2)), dw.copy_(dev(w2))
        g.launch()
        _, _, eager = P.mst_solve(n, dev(u2), dev(v2), dev(w2), P.W_F32, light_factor=lf)
        ids, lab, st = g.result()
        k_ids, k_lab = O.kruskal(n, u2, v2, w2, 1)
        assert np.array_equal(host(ids), k_ids), f"replay {rep}"
        assert np.array_equal(host(lab), k_lab), f"replay {rep}"
        for k in ("mst_count", "weight_fx", "rounds", "m_log", "c_log"):
            assert st[k] == eager[k], (rep, k)
    g.close()


@pytest.mark.parametrize("lf", [None, 0.0])
def test_graph_too_few_rounds_reports_noconv(P, lf):
    """One round per phase cannot finish a path of 2^12 vertices: the stats
    call reports PHEM_ENOCONV; the graph stays usable and a graph recorded
    with enough rounds solves it."""
    n = 1 << 12
    u, v, w = S.path_chords(n, 1 << 10, 3)
    du, dv, dw = dev(u), dev(v), dev(w)
    g = P.MstGraph(n, du, dv, dw, P.W_F32, rounds=1, light_factor=lf)
    g.launch()
    with pytest.raises(P.PhemError) as e:
        g.stats()
    assert e.value.code == P.ENOCONV
    g.close()
    g = P.MstGraph(n, du, dv, dw, P.W_F32, rounds=0, light_factor=lf)
    g.launch()
    ids, lab, st = g.result()
    k_ids, k_lab = O.kruskal(n, u, v, w, 1)
    assert np.array_equal(host(ids), k_ids) and np.array_equal(host(lab), k_lab)
    g.close()


def test_graph_rejects_bad_arguments(P):
    u = dev(np.array([0, 1], np.uint32))
    with pytest.raises(P.PhemError):
        P.MstGraph(4, u, u, u, wmode=7, rounds=3)
    with pytest.raises(P.PhemError):
        P.MstGraph(4, u, u, u, rounds=-1)
    with pytest.raises(P.PhemError):
        P.MstGraph(4, u, u, u, rounds=64)  # round 0 + 64 > 64 rounds


@pytest.mark.parametrize("lf", [None, 0.0])
@pytest.mark.parametrize("name", ["rmat_small", "grid_small", "u32_weights", "many_components"])
def test_graph_maximum_forest(P, name, lf):
    """The recorded solve with PHEM_MST_MAXIMUM equals the eager maximum
    spanning forest, replay after replay."""
    n, u, v, w, wmode = cases.make(name)
    du, dv, dw = dev(u), dev(v), dev(w)
    e_ids, e_lab, eager = P.mst_solve(n, du, dv, dw, wmode, light_factor=lf, maximum=True)
    e_ids, e_lab = host(e_ids).copy(), host(e_lab).copy()
    g = P.MstGraph(n, du, dv, dw, wmode, rounds=0, light_factor=lf, maximum=True)
    for rep in range(2):
        g.launch()
        ids, lab, st = g.result()
        assert np.array_equal(host(ids), e_ids), rep
        assert np.array_equal(host(lab), e_lab), rep
        for k in LOG_KEYS:
            assert st[k] == eager[k], (rep, k)
        torch.cuda.synchronize()
    g.close()


@pytest.mark.parametrize("name", ["rmat_small", "grid_small", "path_chords_small",
                                  "many_components", "u32_weights", "equal_weight_path"])
def test_round_by_round_state_equals_oracle(P, name):
    """The canonical schedule stopped after r rounds (round 0 plus r - 1
    recorded rounds): the forest edges hooked so far and the min-id labels of
    the components so far equal the oracle's round loop stopped after r
    rounds, for every r; only the last r reports convergence."""
    n, u, v, w, wmode = cases.make(name)
    du, dv, dw = dev(u), dev(v), dev(w)
    full = O.boruvka(n, u, v, w, wmode)
    print(f"case {name}: {full['rounds']} rounds", flush=True)
    for r in range(2, full["rounds"] + 1):
        g = P.MstGraph(n, du, dv, dw, wmode, rounds=r - 1, light_factor=0.0)
        g.launch()
        st = P.MstStats()
        rc = P.lib().phem_mst_graph_stats(g.h, ctypes.byref(st), P._stream())
        assert rc == (0 if r == full["rounds"] else P.ENOCONV), r
        o_ids, o_lab = O.boruvka_rounds(n, u, v, w, wmode, r)
        assert np.array_equal(host(g.ids[: st.mst_count]), o_ids), r
        assert np.array_equal(host(g.labels[:n]), o_lab), r
        g.close()
