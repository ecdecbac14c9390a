"""Seeded graph cases shared by the oracle tests and the GPU parity tests.

Each case is (name, n, u, v, w, wmode) with numpy uint32 arrays; w holds
float32 bit patterns (wmode 1) or unsigned integers (wmode 0).
"""
import numpy as np

import synth as S


def f32(x):
    return np.asarray(x, dtype=np.float32).view(np.uint32)


def _empty():
    z = np.zeros(0, np.uint32)
    return 5, z, z.copy(), z.copy(), 1


def _single_vertex():
    return 1, np.array([0], np.uint32), np.array([0], np.uint32), f32([0.5]), 1


def _self_loops_only():
    u = np.arange(10, dtype=np.uint32)
    return 10, u, u.copy(), f32(np.linspace(0, 1, 10)), 1


def _out_of_range():
    u = np.array([0, 1, 2, 7, 3, 9], np.uint32)
    v = np.array([1, 2, 8, 3, 0xFFFFFFFF, 4], np.uint32)
    return 8, u, v, f32([0.5, 0.25, 0.1, 0.3, 0.0, 0.7]), 1


def _duplicates_equal_w():
    # many parallel edges with identical weight: ties broken by id
    rng = np.random.default_rng(7)
    u = rng.integers(0, 12, 400).astype(np.uint32)
    v = rng.integers(0, 12, 400).astype(np.uint32)
    return 12, u, v, f32(np.full(400, 0.5)), 1


def _equal_weight_path():
    # every vertex's minimum edge points left: one hook chain of length n - 1
    n = 5000
    u = np.arange(n - 1, dtype=np.uint32)
    return n, u, u + 1, f32(np.ones(n - 1)), 1


def _reverse_path():
    n = 5000
    u = np.arange(n - 1, dtype=np.uint32)
    w = f32(np.arange(n - 1, 0, -1, dtype=np.float32))
    return n, u, u + 1, w, 1


def _star():
    n = 20000
    v = np.arange(1, n, dtype=np.uint32)
    u = np.zeros(n - 1, np.uint32)
    rng = np.random.default_rng(3)
    return n, u, v, f32(rng.random(n - 1, dtype=np.float32)), 1


def _star_hub_last():
    # the hub has the largest id: each leaf smaller than the hub's current
    # root hooks that root, which changes many times as the leaves join
    n = 1 << 16
    u = np.arange(n - 1, dtype=np.uint32)
    v = np.full(n - 1, n - 1, np.uint32)
    rng = np.random.default_rng(5)
    return n, u, v, f32(rng.random(n - 1, dtype=np.float32)), 1


def _u32_weights():
    rng = np.random.default_rng(11)
    n, m = 3000, 40000
    u = rng.integers(0, n, m).astype(np.uint32)
    v = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(0, 2**32, m, dtype=np.uint64).astype(np.uint32)
    w[:100] = 0xFFFFFFFF
    w[100:200] = 0
    return n, u, v, w, 0


def _signed_floats():
    rng = np.random.default_rng(5)
    n, m = 2000, 20000
    u = rng.integers(0, n, m).astype(np.uint32)
    v = rng.integers(0, n, m).astype(np.uint32)
    w = rng.standard_normal(m).astype(np.float32)
    w[:50] = 0.0
    w[50:100] = -0.0
    return n, u, v, w.view(np.uint32), 1


def _uniform_c1():
    # plain CPU config: n = 2^16, m = 2^20 uniform, weights in [0, 1)
    u, v, w = S.uniform(1 << 16, 1 << 20, 1)
    return 1 << 16, u, v, w, 1


def _rmat_small():
    u, v, w = S.rmat(16, 1 << 20, 2, positive=True)
    return 1 << 16, u, v, w, 1


def _grid_small():
    u, v, w = S.grid(256, 3)
    return 256 * 256, u, v, w, 1


def _path_chords_small():
    n = 1 << 16
    u, v, w = S.path_chords(n, 1 << 14, 4)
    return n, u, v, w, 1


def _many_components():
    # a sparse uniform graph: many components and isolated vertices
    u, v, w = S.uniform(1 << 15, 1 << 13, 9)
    return 1 << 15, u, v, w, 1


CASES = {
    "empty": _empty,
    "single_vertex": _single_vertex,
    "self_loops_only": _self_loops_only,
    "out_of_range": _out_of_range,
    "duplicates_equal_w": _duplicates_equal_w,
    "equal_weight_path": _equal_weight_path,
    "reverse_path": _reverse_path,
    "star": _star,
    "star_hub_last": _star_hub_last,
    "u32_weights": _u32_weights,
    "signed_floats": _signed_floats,
    "many_components": _many_components,
    "uniform_c1": _uniform_c1,
    "rmat_small": _rmat_small,
    "grid_small": _grid_small,
    "path_chords_small": _path_chords_small,
}


def make(name):
    n, u, v, w, wmode = CASES[name]()
    return (n, np.ascontiguousarray(u, np.uint32), np.ascontiguousarray(v, np.uint32),
            np.ascontiguousarray(w, np.uint32), wmode)
