"""The library's host backend: its minimum spanning forest (phem_cpu_msf, also
the CPU pre-filter of the heterogeneous chunked solve) equals the oracle's
Kruskal forest, and its connected components (phem_cpu_cc) equal BFS labels."""
import numpy as np
import pytest

import cases
import oracle as O
import phem_native as P
import synth as S


@pytest.mark.parametrize("name", list(cases.CASES))
def test_cpu_msf_cases(name):
    n, u, v, w, wmode = cases.make(name)
    k_ids, _ = O.kruskal(n, u, v, w, wmode)
    for threads in (1, 4):
        assert np.array_equal(P.cpu_msf(n, u, v, w, wmode, threads), k_ids)


@pytest.mark.parametrize("wq", [0, 3, 1])
def test_cpu_msf_dense_filter_path(wq):
    """m > 5n/2: the threshold split and the filter of the heavy edges run.
    wq > 0 quantizes the weights to wq distinct integers (ties straddle the
    threshold; wq = 1 makes every edge heavy), 0 keeps float weights."""
    u, v, w = S.rmat(13, 1 << 18, 11)
    n = 1 << 13
    wmode = P.W_F32
    if wq:
        w = np.random.default_rng(5).integers(0, wq, len(u)).astype(np.uint32)
        wmode = P.W_U32
    want = O.kruskal(n, u, v, w, wmode)[0]
    for threads in (1, 4, 16):
        assert np.array_equal(P.cpu_msf(n, u, v, w, wmode, threads), want)


def test_cpu_msf_tie_heavy_integer_weights():
    u, v, w = S.grid(300, 7)
    w = (w.view(np.float32)).astype(np.uint32)  # integers 1..16
    n = 300 * 300
    assert np.array_equal(P.cpu_msf(n, u, v, w, P.W_U32, 8), O.kruskal(n, u, v, w, 0)[0])


@pytest.mark.parametrize("name", list(cases.CASES))
def test_cpu_cc_cases(name):
    n, u, v, w, wmode = cases.make(name)
    want = O.bfs_labels(n, u, v)
    for threads in (1, 8):
        assert np.array_equal(P.cpu_cc(n, u, v, threads), want)


def test_cpu_cc_rmat_threads():
    u, v, _ = S.rmat(16, 1 << 20, 5)
    n = 1 << 16
    assert np.array_equal(P.cpu_cc(n, u, v, 8), O.uf_labels(n, u, v))
