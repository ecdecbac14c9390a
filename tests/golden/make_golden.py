"""Writes tests/golden/fixtures.json: for every seeded case of tests/cases.py
and a few larger synthetic graphs, digests of the input and the oracle's
answers (forest ids, exact weight, min-id labels, the canonical Boruvka
round log, the maximum spanning forest).  The reference has no golden
vectors of its own for this path, so these pin the build's restatement
against drift.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import cases  # noqa: E402
import max_forest as MF  # noqa: E402
import oracle as O  # noqa: E402
import synth as S  # noqa: E402


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint32).tobytes()).hexdigest()


# name -> (n, u, v, w, wmode), beyond tests/cases.py
EXTRA = {
    "rmat_14_18_seed3": lambda: (1 << 14,) + S.rmat(14, 1 << 18, 3, positive=True) + (1,),
    "uniform_14_17_seed4": lambda: (1 << 14,) + S.uniform(1 << 14, 1 << 17, 4) + (1,),
    "grid_128_seed5": lambda: (128 * 128,) + S.grid(128, 5) + (1,),
    "path_chords_14_seed6": lambda: (1 << 14,) + S.path_chords(1 << 14, 1 << 12, 6) + (1,),
}


def make(name):
    if name in cases.CASES:
        return cases.make(name)
    n, u, v, w, wmode = EXTRA[name]()
    return (n, np.ascontiguousarray(u, np.uint32), np.ascontiguousarray(v, np.uint32),
            np.ascontiguousarray(w, np.uint32), wmode)


def names():
    return list(cases.CASES) + list(EXTRA)


def expected(name):
    n, u, v, w, wmode = make(name)
    ids, lab = O.kruskal(n, u, v, w, wmode)
    ws, exact = O.weight_sum(ids, w, wmode)
    b = O.boruvka(n, u, v, w, wmode)
    x_ids, _ = MF.max_forest_oracle(n, u, v, w, wmode)
    return {
        "n": int(n), "m": int(len(u)), "wmode": int(wmode),
        "input_sha256": hashlib.sha256(u.tobytes() + v.tobytes() + w.tobytes()).hexdigest(),
        "forest_count": int(len(ids)), "forest_sha256": digest(ids),
        "weight_fx": str(ws), "weight_exact": bool(exact),
        "labels_sha256": digest(lab), "components": int(len(np.unique(lab))) if n else 0,
        "round_log_m": [int(x) for x in b["log_m"]],
        "round_log_c": [int(x) for x in b["log_c"]],
        "max_forest_sha256": digest(x_ids),
    }


def main():
    out = {name: expected(name) for name in names()}
    path = os.path.join(HERE, "fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {len(out)} fixtures to {path}")


if __name__ == "__main__":
    main()
