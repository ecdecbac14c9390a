"""The host minimum spanning forest (phem_cpu_msf, the CPU pre-filter of the
heterogeneous chunked solve) equals the oracle's Kruskal forest."""
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


def test_cpu_msf_tie_heavy_integer_weights():
    u, v, w = S.grid(300, 7)
    w = (w.view(np.float32)).astype(np.uint32)  # integers 1..16
    n = 300 * 300
    assert np.array_equal(P.cpu_msf(n, u, v, w, P.W_U32, 8), O.kruskal(n, u, v, w, 0)[0])
