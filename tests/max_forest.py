"""Maximum spanning forest expected by the tests, from the oracle."""
import numpy as np

import oracle as O


def max_forest_oracle(n, u, v, w, wmode):
    """Maximum spanning forest under (weight, id) descending, by reduction to
    the oracle's minimum forest: reversing the edge order and the order key
    turns (w, id) descending into (key', id') ascending; the reversed keys go
    in as unsigned weights, whose order key is the identity."""
    m = len(u)
    kr = (~O.wkey(w, wmode)[::-1]).astype(np.uint32)
    ids, lab = O.kruskal(n, u[::-1].copy(), v[::-1].copy(), kr, 0)
    return np.sort((m - 1 - ids.astype(np.int64))).astype(np.uint32), lab
