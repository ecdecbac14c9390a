"""The package's workload definitions agree with the test-side restatement."""
import numpy as np

import synth as S
import phem_workloads as W


def test_mix_params_match_restatement():
    for g in range(3000):
        assert W.mix_params(5, g) == S.mix_graph_params(5, g), g


def test_mix_distribution():
    ps = [W.mix_params(1, g) for g in range(4096)]
    n = np.array([p[0] for p in ps], dtype=np.float64)
    m = np.array([p[1] for p in ps], dtype=np.float64)
    assert n.min() >= 256 and n.max() <= 1 << 20
    d = 2 * m / n
    assert 3.9 <= d.min() and d.max() <= 32.1
    assert abs(np.log2(n).mean() - 14) < 0.3  # log-uniform in [8, 20]
    assert [p[2] for p in ps[:4]] == [0, 1, 0, 1]


def test_stream_seed_matches():
    for b in range(10):
        assert W.stream_seed(4, b) == S.stream_seed(4, b)


def test_draw_matches_numpy():
    for i in [0, 1, 12345, 2**40 + 7]:
        assert W.draw(9, 3, i) == int(S.draw(9, 3, np.uint64(i)))
