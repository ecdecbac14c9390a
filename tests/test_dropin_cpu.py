"""The C++ drop-in new_boruvka in its CPU-only mode (no GPU work): the forest
and total weight equal the oracle's Kruskal forest, with the library's host
solver and with a caller's solver registered through phem_set_cpu_mst
(MST/src/new_boruvka.cu:468-492)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "harness", "libdropin_mst.so")

pytestmark = pytest.mark.skipif(not os.path.exists(HARNESS), reason="harness not built")

# run in a child process: the harness prints through the C runtime, and the
# printed total weight is part of what is checked
CHILD = r"""
import ctypes, sys
L = ctypes.CDLL(sys.argv[1])
L.harness_new_boruvka.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
L.harness_register_cpu_mst(int(sys.argv[5]))
L.harness_new_boruvka(sys.argv[2].encode(), sys.argv[3].encode(), 0, 0, int(sys.argv[4]),
                      4, 0)
sys.stdout.flush()
print("calls", L.harness_cpu_mst_calls(), flush=True)
"""


@pytest.mark.parametrize("registered", [0, 1])
@pytest.mark.parametrize("gen", [0, 1])
def test_new_boruvka_cpu_only(tmp_path, gen, registered):
    rng = np.random.default_rng(7 + gen)
    n, m = 2000, 30000
    a = rng.integers(1, n + 1, m)
    b = rng.integers(1, n + 1, m)
    w = rng.integers(0, 40, m)
    lines = [f"{x} {y}" + ("" if gen else f" {z}") for x, y, z in zip(a, b, w)]
    path = tmp_path / "g.mtx"
    path.write_text("%%MatrixMarket matrix coordinate integer general\n"
                    f"{n} {n} {m}\n" + "\n".join(lines) + "\n")
    res = tmp_path / "forest.txt"
    out = subprocess.run([sys.executable, "-c", CHILD, HARNESS, str(path), str(res),
                          str(gen), str(registered)],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    import phem_native as P
    g = P.read_mtx(str(path), generate_weights=bool(gen))
    ids, _ = O.kruskal(g["n"], g["u"], g["v"], g["w"], 0)
    want = np.stack([g["u"][ids] + 1, g["v"][ids] + 1, g["w"][ids]], axis=1).astype(np.uint64)
    got = np.loadtxt(res, dtype=np.uint64, ndmin=2)
    assert got.shape == want.shape and np.array_equal(got, want)
    total = re.search(r"Total weight (\d+)", out.stdout)
    assert total and int(total.group(1)) == int(g["w"][ids].astype(np.uint64).sum())
    assert re.search(r"pbbs execution time \d+ microseconds", out.stdout)
    assert f"calls {registered}" in out.stdout
