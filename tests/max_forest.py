"""Maximum spanning forest expected by the tests, from the oracle."""
import numpy as np

import oracle as O


def max_forest_oracle(n, u, v, w, wmode):
    """Maximum spanning forest under (weight, id) descending, by reduction to
    the oracle's minimum forest: reversing the edge order and the weight
    order turns (w, id) descending into (w', id') ascending."""
    m = len(u)
    kr = ~O.wkey(w, wmode)[::-1]
    if wmode == 0:
        wr = kr
    else:  # float bit patterns whose order key is kr
        wr = np.where(kr & np.uint32(0x80000000), kr & np.uint32(0x7FFFFFFF), ~kr)
    ids, lab = O.kruskal(n, u[::-1].copy(), v[::-1].copy(), wr.astype(np.uint32), wmode)
    return np.sort((m - 1 - ids.astype(np.int64))).astype(np.uint32), lab

