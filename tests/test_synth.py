"""Synthetic generator definitions (CPU): structural and statistical
properties of the byte streams the device generators must reproduce."""
import numpy as np

import synth as S


def test_splitmix64_known_values():
    # splitmix64 of state 0 (first output of the published generator seeded 0)
    assert int(S.mix64(np.uint64(0))) == 0xE220A8397B1DCDAF
    assert int(S.mix64(np.uint64(0x9E3779B97F4A7C15))) == 0x6E789E6AA1B965F4


def test_uniform_structure():
    n, m = 1000, 50000
    u, v, w = S.uniform(n, m, 42)
    assert (u < v).all() and (v < n).all()
    f = w.view(np.float32)
    assert (f >= 0).all() and (f < 1).all()
    k = f.astype(np.float64) * 2.0 ** 24
    assert (k == np.floor(k)).all()
    # endpoints roughly uniform: each decile of ids holds ~20% of endpoints
    h = np.bincount(np.concatenate([u, v]) * 10 // n, minlength=10) / (2 * m)
    assert np.all(np.abs(h - 0.1) < 0.01)
    u2, v2, w2 = S.uniform(n, m, 42)
    assert np.array_equal(u, u2) and np.array_equal(w, w2)
    u3, _, _ = S.uniform(n, m, 43)
    assert not np.array_equal(u, u3)


def test_uniform_id0_is_a_window():
    u, v, w = S.uniform(500, 3000, 7)
    u2, v2, w2 = S.uniform(500, 1000, 7, id0=2000)
    assert np.array_equal(u[2000:], u2) and np.array_equal(w[2000:], w2)


def test_positive_weights():
    _, _, w = S.uniform(100, 20000, 3, positive=True)
    f = w.view(np.float32)
    assert (f > 0).all() and (f <= 1).all()


def test_rmat_quadrant_frequencies():
    scale, m = 12, 1 << 16
    u, v, w = S.rmat(scale, m, 2)
    assert (u < v).all() and (v < (1 << scale)).all()
    # top-level bit of the (unordered) pair: P(both 0) ~ a = 0.57 and
    # P(both 1) ~ d = 0.05 (conditioned on no self-loop, approximately)
    top = scale - 1
    bu, bv = (u >> top) & 1, (v >> top) & 1
    assert abs(np.mean((bu == 0) & (bv == 0)) - 0.57) < 0.02
    assert abs(np.mean((bu == 1) & (bv == 1)) - 0.05) < 0.01
    assert S.rmat_thresholds(0.57, 0.19, 0.19) == S.RMAT_T


def test_grid_structure():
    W = 64
    u, v, w = S.grid(W, 3)
    H = W * (W - 1)
    assert abs(len(u) / (2 * H) - 0.8) < 0.03
    d = v.astype(np.int64) - u.astype(np.int64)
    assert set(np.unique(d)) <= {1, W}
    assert ((u % W != W - 1) | (d == W)).all()   # no wrap-around
    f = w.view(np.float32)
    assert set(np.unique(f)) <= set(range(1, 17))


def test_path_chords_structure():
    n, k = 4096, 512
    u, v, w = S.path_chords(n, k, 4)
    assert len(u) == n - 1 + k
    deg = np.bincount(np.concatenate([u[: n - 1], v[: n - 1]]), minlength=n)
    assert sorted(np.bincount(deg).nonzero()[0].tolist()) == [1, 2]
    assert (deg == 1).sum() == 2     # a Hamiltonian path


def test_mix_params_ranges():
    for g in range(200):
        n, m, kind, seed = S.mix_graph_params(5, g)
        assert 256 <= n <= (1 << 20) + 1
        assert 2 * n <= m <= 16 * n
        if kind == 1:
            assert n & (n - 1) == 0
