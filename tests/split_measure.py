"""Measures the CPU/GPU splits of the two heterogeneous paths on this host and
its GPU, to choose their defaults:

- MST: phem_msf_host_chunks (host edge lists in chunks) with the host
  reducing the first cpu_edges edges of the last chunk (the reference's
  heterogeneous streamline, MST/src/kernels.cu:1155-1826, which gives the
  host min(last chunk, 10^8) edges);
- CC: the reference's HAST driver flow (CC/main.cu:43-203) with the
  reference's own host_spanning_tree on the host share (oracle/_ref build)
  over the library's SpanningTree, at GPU_share values (CC/main.cu:59-65).

Workloads: RMAT n=2^24, m=2^28 (seed 1) and the streaming list (RMAT
n=2^22, 2^24 edges, seed 4, then 256 uniform batches of 2^20), both
generated on the device and copied to host memory.  Every run's result is
compared with the first run of its workload (forest ids, or labels).
Test infrastructure: run by hand, `python tests/split_measure.py [mst]
[cc]`; prints one line per run and a JSON summary as its last line.
"""
import ctypes
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "phem-graph_amd"))

import numpy as np  # noqa: E402
import torch  # noqa: E402

import phem_native as P  # noqa: E402
import phem_workloads as W  # noqa: E402

HAST = os.path.join(ROOT, "oracle", "_ref", "libref_hast.so")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def log(*a):
    print(*a, flush=True)


def host(t):
    return t.cpu().numpy().view(np.uint32)


def workloads(which):
    dev = torch.device("cuda:0")
    if which == "rmat24":
        g = P.gen_rmat(24, 1 << 28, 1, device=dev)
        out = [host(x) for x in g]
        return 1 << 24, out, 1 << 26
    n = 1 << 22
    parts = [[host(x) for x in P.gen_rmat(22, 1 << 24, 4, device=dev)]]
    for b in range(256):
        parts.append([host(x) for x in P.gen_uniform(n, 1 << 20, W.stream_seed(4, b),
                                                      device=dev)])
    out = [np.concatenate([p[i] for p in parts]) for i in range(3)]
    return n, out, 1 << 24


def digest(a):
    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def measure_mst(name, n, uvw, chunk, reps=2):
    u, v, w = uvw
    m = len(u)
    last = ((m + chunk - 1) // chunk - 1) * chunk
    shares = [0, 1 << 20, 1 << 22, 1 << 24, min(m - last, 1 << 26), min(m - last, 10 ** 8)]
    rows, ref = [], None
    for cpu in sorted(set(shares)):
        best = None
        for _ in range(reps):
            ids, st = P.msf_host_chunks(n, u, v, w, P.W_F32, chunk, cpu)
            d = digest(ids)
            ref = ref or d
            assert d == ref, f"{name}: forest differs at cpu_edges={cpu}"
            if best is None or st["seconds"] < best["seconds"]:
                best = st
        row = {"cpu_edges": best["cpu_edges"], "cpu_kept": best["cpu_kept"],
               "s": round(best["seconds"], 4), "cpu_s": round(best["cpu_seconds"], 4),
               "edges_per_s": m / best["seconds"]}
        log(f"mst {name} chunk={chunk} cpu_edges={cpu}: {row}")
        rows.append(row)
    # rates: the device's over the chunks before the last, the host's share
    base = rows[0]["s"]
    cpu_rows = [r for r in rows if r["cpu_edges"]]
    host_rate = max(r["cpu_edges"] / r["cpu_s"] for r in cpu_rows) if cpu_rows else None
    return {"workload": name, "n": n, "m": m, "chunk": chunk, "last_chunk_first": last,
            "runs": rows, "device_edges_per_s": m / base,
            "host_edges_per_s": host_rate,
            "best_cpu_edges": min(rows, key=lambda r: r["s"])["cpu_edges"]}


def measure_cc(name, n, uvw, batch=1 << 22, reps=2):
    L = ctypes.CDLL(HAST)
    L.ref_hast_run.restype = ctypes.c_int
    L.ref_hast_run.argtypes = [ctypes.c_int, ctypes.c_long, _u64p, ctypes.c_float,
                               ctypes.c_int, _u32p, _u64p, _u64p,
                               ctypes.POINTER(ctypes.c_float)]
    u, v, _ = uvw
    rec = ((np.minimum(u, v).astype(np.uint64) << np.uint64(32))
           | np.maximum(u, v).astype(np.uint64))
    m = len(rec)
    lab = np.zeros(n, np.uint32)
    t1 = np.zeros(n, np.uint64)
    t2 = np.zeros(n, np.uint64)
    rows, ref = [], None
    for share in (1.0, 0.95, 0.9, 0.85, 0.8, 0.75, 0.7, 0.6, 0.5):
        best = None
        for _ in range(reps):
            ms = ctypes.c_float(0)
            assert L.ref_hast_run(n, m, rec, share, batch, lab, t1, t2, ctypes.byref(ms)) == 0
            d = digest(lab)
            ref = ref or d
            assert d == ref, f"{name}: labels differ at GPU_share={share}"
            best = ms.value if best is None else min(best, ms.value)
        row = {"gpu_share": share, "ms": round(best, 3), "edges_per_s": m / (best * 1e-3)}
        log(f"cc {name} batch={batch}: {row}")
        rows.append(row)
    return {"workload": name, "n": n, "m": m, "batch": batch, "runs": rows,
            "best_gpu_share": min(rows, key=lambda r: r["ms"])["gpu_share"],
            "host_threads": os.environ.get("PARLAY_NUM_THREADS")}


def main():
    """Arguments: any of `mst`, `cc` (default both)."""
    parts = set(sys.argv[1:]) or {"mst", "cc"}
    P.require_device()
    out = {"host_cpus": len(os.sched_getaffinity(0))}
    for name in ("rmat24", "stream22"):
        t = time.time()
        n, uvw, chunk = workloads(name)
        log(f"{name}: n={n} m={len(uvw[0])} generated in {time.time() - t:.1f} s")
        if "mst" in parts:
            out[f"mst_{name}"] = measure_mst(name, n, uvw, chunk)
        if "cc" not in parts:
            pass
        elif os.path.exists(HAST):
            out[f"cc_{name}"] = measure_cc(name, n, uvw)
        else:
            log("oracle/_ref/libref_hast.so missing: CC split not measured")
        del uvw
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
