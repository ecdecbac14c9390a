"""CPU checks of bench.py's contract pieces that need no GPU: the multi-GPU
guard exits with a message, and the algorithmic-byte formulas follow their
definitions (B_MST = 32 S_m - 20 m + 16 S_c + 12 n over the canonical round
log, which the oracle restates; B_CC = 8 m + 12 n)."""
import os
import subprocess
import sys

import numpy as np

import cases
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_b_mst_from_the_canonical_round_log():
    n, u, v, w, wmode = cases.make("rmat_small")
    b = O.boruvka(n, u, v, w, wmode)
    lm, lc = [int(x) for x in b["log_m"]], [int(x) for x in b["log_c"]]
    m = len(u)
    # round 0 reads 12 B per input edge; later rounds read 16 B per alive
    # edge; every round writes 16 B per survivor; 16 B per active component
    # per round; 12 B per vertex
    reads = 12 * m + sum(16 * x for x in lm[1:])
    writes = sum(16 * x for x in lm[1:])
    direct = reads + writes + 16 * sum(lc) + 12 * n
    # the closed form counts round 0 as m alive edges (S_m includes m)
    S_m = m + sum(lm[1:])
    assert bench.b_mst(m, n, [m] + lm[1:], lc) == 32 * S_m - 20 * m + 16 * sum(lc) + 12 * n
    assert bench.b_mst(m, n, [m] + lm[1:], lc) == direct
    assert bench.b_cc(m, n) == 8 * m + 12 * n


def test_kernel_bytes_of_a_split_log():
    st = {"m_log": [100, 60, 20, 30, 5], "light_rounds": 3, "tau": 7,
          "light_edges": 100, "heavy_edges": 30, "giant_edges": 10, "c_log": [50, 20, 5, 12, 3]}
    m, n = 1000, 80
    kb = bench.mst_kernel_bytes(st, m, n)
    # survivors: light rounds 0->1 (60), 1->2 (20), none after the light
    # phase's last round; heavy round 3->4 (5)
    assert kb["k_compact"] == 16 * (100 + 60 + 20 + 30 + 5) + 16 * (60 + 20 + 0 + 5 + 0)
    assert kb["k_light"] == 12 * m + 16 * 100
    assert kb["k_heavy"] == 12 * m + 16 * (30 - 10)
    assert kb["k_giant_emit"] == 8 * n + 16 * 10
    assert bench.b_actual(st, m, n) == sum(kb.values()) - kb["k_labels"] + 16 * sum(st["c_log"]) + 12 * n
