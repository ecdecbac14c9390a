"""Synthetic workloads: the serving mix of independent graphs and the
streaming insertion batches, built on the device generators of libphem.so.

Serving mix: graph g has n = floor(2^(8 + 12 U)) vertices (U uniform in
[0, 1), RMAT graphs rounded to a power of two), average degree d uniform in
4..32 and m = n d / 2 edges; even g are uniform random graphs and odd g RMAT
(a, b, c, d) = (.57, .19, .19, .05), float32 weights uniform in [0, 1).
All draws come from the same counter-based splitmix64 streams as the device
generators (draw(seed, stream, i) = mix64(mix64(seed * GOLDEN + stream) + i)).
"""
import math

import phem_native as P

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def mix64(x):
    z = (x + GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def draw(seed, stream, i):
    return mix64((mix64((seed * GOLDEN + stream) & M64) + i) & M64)


def mulhi(r, n):
    return ((r >> 32) * n) >> 32


def mix_params(seed, g):
    """(n, m, kind, graph seed) of graph g of the serving mix (kind 0 uniform,
    1 RMAT)."""
    U = (draw(seed, 10, g) >> 11) * 2.0 ** -53
    n = int(math.floor(2.0 ** (8 + 12 * U)))
    d = 4 + mulhi(draw(seed, 11, g), 29)
    kind = g % 2
    if kind == 1:
        n = 1 << int(round(math.log2(n)))
    return n, n * d // 2, kind, (seed * 7919 + g) & M64


def mix_graph(seed, g, device="cuda", out=None):
    """Graph g of the serving mix on the device: (n, u, v, w) int32 tensors."""
    n, m, kind, gs = mix_params(seed, g)
    if kind == 0:
        u, v, w = P.gen_uniform(n, m, gs, device=device, out=out)
    else:
        u, v, w = P.gen_rmat(n.bit_length() - 1, m, gs, device=device, out=out)
    return n, u, v, w


def stream_seed(seed, b):
    """Seed of streaming insertion batch b."""
    return (seed + 0x1000 * (b + 1)) & M64
