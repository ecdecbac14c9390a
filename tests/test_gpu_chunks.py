"""GPU parity of the chunked solve over host-resident edges
(phem_msf_host_chunks): for every chunking and CPU share the forest ids equal
the oracle's Kruskal forest of the whole edge list under the (weight, id)
order, and the exact weight agrees."""
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


@pytest.mark.parametrize("name", list(cases.CASES))
def test_chunks_cases(P, name):
    n, u, v, w, wmode = cases.make(name)
    want = O.kruskal(n, u, v, w, wmode)[0]
    m = len(u)
    # one chunk, three, and up to 64 (7 edges each for the small cases)
    for chunk in sorted({max(1, m), max(1, m // 3), max(7, m // 64)}):
        for cpu in (0, 5, P.CPU_EDGES_AUTO, 1 << 40):
            print(f"case {name} m={m} chunk={chunk} cpu={cpu}", flush=True)
            ids, st = P.msf_host_chunks(n, u, v, w, wmode, chunk, cpu)
            assert np.array_equal(ids, want), (chunk, cpu)
            assert st["chunks"] == (m + chunk - 1) // chunk


@pytest.mark.parametrize("wq", [0, 3])
def test_chunks_dense_cpu_share(P, wq):
    """RMAT 2^14 with 2^20 edges (the host share is dense enough for its
    forest to drop most of it), float weights or 3 integer weights (ties
    across chunk borders and across the host share)."""
    u, v, w = S.rmat(14, 1 << 20, 21)
    n = 1 << 14
    wmode = P.W_F32
    if wq:
        w = np.random.default_rng(2).integers(0, wq, len(u)).astype(np.uint32)
        wmode = P.W_U32
    want = O.kruskal(n, u, v, w, wmode)[0]
    ws, _ = O.weight_sum(want, w, wmode)
    for chunk, cpu in ((1 << 18, 0), (1 << 18, 1 << 17), (1 << 18, 1 << 18),
                       (300000, P.CPU_EDGES_AUTO), ((1 << 20) - 1, 1 << 30)):
        print(f"dense wq={wq} chunk={chunk} cpu={cpu}", flush=True)
        ids, st = P.msf_host_chunks(n, u, v, w, wmode, chunk, cpu)
        assert np.array_equal(ids, want), (chunk, cpu)
        assert O.weight_sum(ids, w, wmode)[0] == ws
        last = (st["chunks"] - 1) * chunk
        if cpu and st["chunks"] > 1 and cpu != P.CPU_EDGES_AUTO:
            assert st["cpu_edges"] == min(cpu, len(u) - last)
        assert st["cpu_kept"] <= st["cpu_edges"]
        if st["cpu_edges"] >= 1 << 17:
            assert st["cpu_kept"] < n  # a forest of the share


def test_chunks_auto_rule(P):
    """PHEM_CPU_EDGES_AUTO gives the host cpu_share x the edges before the
    last chunk (capped at the last chunk)."""
    u, v, w = S.rmat(12, 1 << 18, 4)
    n = 1 << 12
    want = O.kruskal(n, u, v, w, 1)[0]
    old = P.get_tuning()["cpu_share"]
    try:
        for share, expect in ((0.0, 0), (0.1, int(0.1 * 3 * 65536)), (5.0, 65536)):
            P.set_tuning(cpu_share=share)
            ids, st = P.msf_host_chunks(n, u, v, w, P.W_F32, 65536, P.CPU_EDGES_AUTO)
            assert np.array_equal(ids, want)
            assert st["cpu_edges"] == expect
    finally:
        P.set_tuning(cpu_share=old)


def test_chunks_refusals(P):
    u = v = w = np.zeros(4, np.uint32)
    for args in ((0, 4, 1, 0), (8, 4, 0, 0)):
        n, m, chunk, cpu = args
        with pytest.raises(P.PhemError):
            P.msf_host_chunks(n, u[:m], v[:m], w[:m], P.W_U32, chunk, cpu)


@pytest.mark.slow
def test_chunks_rmat_2e22_2e26(P):
    """A C4-scale host list: RMAT 2^22 with 2^26 edges in four chunks of
    2^24, with and without the host share, against the parallel oracle."""
    u, v, w = S.rmat(22, 1 << 26, 3)
    n = 1 << 22
    want = O.par_mst(n, u, v, w, 1)
    for cpu in (0, P.CPU_EDGES_AUTO, 1 << 23):
        print(f"rmat22 chunk=2^24 cpu={cpu}", flush=True)
        ids, st = P.msf_host_chunks(n, u, v, w, P.W_F32, 1 << 24, cpu)
        assert np.array_equal(ids, want)
        print(st, flush=True)
