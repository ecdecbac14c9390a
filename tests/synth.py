"""numpy restatement of the product's synthetic graph generators.

The product generates the benchmark inputs on the device (and the grid and
path generators on the host); this module defines the same byte streams
independently so that tests can (a) feed the CPU oracle without a GPU and
(b) check the product's generators bit for bit.

Counter-based RNG: draw(seed, stream, i) = splitmix64(base(seed, stream) + i),
base(seed, stream) = splitmix64(seed * GOLDEN + stream) (all arithmetic mod 2^64).
Weights are exact float32 values k * 2^-24: k uniform in [0, 2^24) for "[0,1)"
and in [1, 2^24] for "(0,1]".  Every generated edge is stored with u < v.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)
M32 = np.uint64(0xFFFFFFFF)

# RMAT (a, b, c, d) = (0.57, 0.19, 0.19, 0.05) as cumulative 32-bit thresholds:
# t1 = round(a 2^32), t2 = round((a+b) 2^32), t3 = round((a+b+c) 2^32)
RMAT_T = (2448131359, 3264175145, 4080218931)
GRID_KEEP_T = 3435973837  # round(0.8 * 2^32)


def rmat_thresholds(a, b, c):
    return (int(round(a * 2**32)), int(round((a + b) * 2**32)),
            int(round((a + b + c) * 2**32)))


def mix64(x):
    with np.errstate(over="ignore"):
        z = np.asarray(x, dtype=np.uint64) + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
        return z ^ (z >> np.uint64(31))


def stream_base(seed, stream):
    with np.errstate(over="ignore"):
        return mix64(np.uint64(seed) * GOLDEN + np.uint64(stream))


def draw(seed, stream, idx):
    with np.errstate(over="ignore"):
        return mix64(stream_base(seed, stream) + np.asarray(idx, dtype=np.uint64))


def mulhi(r64, n):
    """Map the top 32 bits of r64 to [0, n) by a 32x32 multiply-high."""
    return (((r64 >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)).astype(np.uint32)


def weights(seed, idx, positive=False):
    k = (draw(seed, 2, idx) >> np.uint64(40)).astype(np.uint32)
    if positive:
        k = k + np.uint32(1)
    return (k.astype(np.float32) * np.float32(2.0 ** -24)).view(np.uint32)


def _order(u, v):
    return np.minimum(u, v).astype(np.uint32), np.maximum(u, v).astype(np.uint32)


def uniform(n, m, seed, positive=False, s_u=0, s_v=1, s_re=3, id0=0):
    """Uniform endpoints; self-loops redrawn (64 tries, then v = (u+1) mod n)."""
    i = np.arange(id0, id0 + m, dtype=np.uint64)
    u = mulhi(draw(seed, s_u, i), n)
    v = mulhi(draw(seed, s_v, i), n)
    for t in range(64):
        bad = np.nonzero(u == v)[0]
        if bad.size == 0:
            break
        v[bad] = mulhi(draw(seed, s_re, i[bad] * np.uint64(64) + np.uint64(t)), n)
    bad = u == v
    v[bad] = ((u[bad].astype(np.uint64) + 1) % np.uint64(n)).astype(np.uint32)
    u, v = _order(u, v)
    return u, v, weights(seed, i, positive)


def rmat(scale, m, seed, t=RMAT_T, positive=False):
    """RMAT on 2^scale vertices: one 32-bit draw per level, top level first;
    a self-loop redraws the whole edge (64 tries, then v = u ^ 1)."""
    t1, t2, t3 = (np.uint32(x) for x in t)
    i = np.arange(m, dtype=np.uint64)
    u = np.zeros(m, np.uint32)
    v = np.zeros(m, np.uint32)
    todo = np.arange(m)
    for a in range(64):
        if todo.size == 0:
            break
        ii = i[todo]
        uu = np.zeros(todo.size, np.uint32)
        vv = np.zeros(todo.size, np.uint32)
        for lvl in range(scale):
            h = draw(seed, 8 + a, ii * np.uint64(16) + np.uint64(lvl // 2))
            x = (h >> np.uint64(32)) if lvl % 2 == 0 else (h & M32)
            x = x.astype(np.uint32)
            bu = x >= t2
            bv = ((x >= t1) & (x < t2)) | (x >= t3)
            uu = (uu << np.uint32(1)) | bu.astype(np.uint32)
            vv = (vv << np.uint32(1)) | bv.astype(np.uint32)
        u[todo] = uu
        v[todo] = vv
        todo = todo[uu == vv]
    v[todo] = u[todo] ^ np.uint32(1)
    u, v = _order(u, v)
    return u, v, weights(seed, i, positive)


def grid(W, seed, keep_t=GRID_KEEP_T, wmax=16):
    """W x W 4-neighbour grid: horizontal edges first (row-major), then
    vertical; each kept if draw < keep_t; integer weights 1..wmax as float32.
    Kept edges are numbered consecutively in that order."""
    H = W * (W - 1)
    j = np.arange(2 * H, dtype=np.uint64)
    a = np.empty(2 * H, np.uint64)
    b = np.empty(2 * H, np.uint64)
    jh = j[:H]
    a[:H] = (jh // np.uint64(W - 1)) * np.uint64(W) + jh % np.uint64(W - 1)
    b[:H] = a[:H] + np.uint64(1)
    jv = j[H:] - np.uint64(H)
    a[H:] = jv
    b[H:] = jv + np.uint64(W)
    keep = (draw(seed, 4, j) >> np.uint64(32)) < np.uint64(keep_t)
    w = (mulhi(draw(seed, 5, j), wmax) + np.uint32(1)).astype(np.float32).view(np.uint32)
    return a[keep].astype(np.uint32), b[keep].astype(np.uint32), w[keep]


def path_chords(n, k, seed):
    """Random-permuted path on n vertices (edges 0..n-2: pi[i]-pi[i+1], pi =
    stable argsort of draw(seed, 6, i)) plus k uniform chords (ids n-1..)."""
    pi = np.argsort(draw(seed, 6, np.arange(n, dtype=np.uint64)), kind="stable").astype(np.uint32)
    pu, pv = pi[:-1], pi[1:]
    cu, cv, _ = uniform(n, k, seed, s_u=7, s_v=8, s_re=9)
    u = np.concatenate([pu, cu])
    v = np.concatenate([pv, cv])
    u, v = _order(u, v)
    w = weights(seed, np.arange(len(u), dtype=np.uint64))
    return u, v, w


def stream_seed(seed, b):
    """Seed of streaming insertion batch b."""
    return (seed + 0x1000 * (b + 1)) & 0xFFFFFFFFFFFFFFFF


def mix_graph_params(seed, g):
    """(n, m, kind, graph seed) of graph g of the serving mix: n log-uniform in
    [2^8, 2^20] (RMAT graphs rounded to a power of two), degree d uniform in
    4..32, m = n d / 2; even g uniform, odd g RMAT."""
    r = int(draw(seed, 10, np.uint64(g)))
    U = (r >> 11) * 2.0 ** -53
    n = int(np.floor(2.0 ** (8 + 12 * U)))
    d = 4 + int(mulhi(draw(seed, 11, np.uint64(g)), 29))
    kind = g % 2
    if kind == 1:
        scale = int(round(np.log2(n)))
        n = 1 << scale
    m = n * d // 2
    return n, m, kind, (seed * 7919 + g) & 0xFFFFFFFFFFFFFFFF
