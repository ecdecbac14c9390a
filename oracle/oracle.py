"""ctypes wrapper of the CPU oracle (oracle/libphem_oracle.so) and of the
reference CPU builds under oracle/_ref/.

TEST INFRASTRUCTURE ONLY.  Imported by the test suite, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphem_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")
NO_EDGE = 0xFFFFFFFF

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_lib = None


def build():
    """Compile the oracle (always) and, if the reference tree exists, _ref."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "libphem_oracle.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "-B", "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        c = ctypes
        L.or_kruskal.restype = c.c_int64
        L.or_kruskal.argtypes = [c.c_uint32, c.c_uint64, _u32p, _u32p, _u32p, c.c_int,
                                 _u32p, c.c_void_p]
        L.or_boruvka.restype = c.c_int64
        L.or_boruvka.argtypes = [c.c_uint32, c.c_uint64, _u32p, _u32p, _u32p, c.c_int,
                                 _u32p, c.c_void_p, c.c_void_p, c.c_void_p, c.c_int,
                                 c.POINTER(c.c_int)]
        L.or_boruvka_rounds.restype = c.c_int64
        L.or_boruvka_rounds.argtypes = [c.c_uint32, c.c_uint64, _u32p, _u32p, _u32p,
                                        c.c_int, c.c_int, _u32p, _u32p]
        L.or_uf_labels.restype = None
        L.or_uf_labels.argtypes = [c.c_uint32, c.c_uint64, _u32p, _u32p, _u32p, c.c_void_p]
        L.or_bfs_labels.restype = None
        L.or_bfs_labels.argtypes = [c.c_uint32, c.c_uint64, _u32p, _u32p, _u32p]
        L.or_check_forest.restype = c.c_int
        L.or_check_forest.argtypes = [c.c_uint32, c.c_uint64, _u32p, _u32p, _u32p,
                                      c.c_uint64, _u32p]
        L.or_weight_sum.restype = c.c_uint64
        L.or_weight_sum.argtypes = [c.c_uint64, _u32p, _u32p, c.c_int, c.POINTER(c.c_int)]
        L.or_par_mst.restype = c.c_int64
        L.or_par_mst.argtypes = [c.c_uint32, c.c_uint64, _u32p, _u32p, _u32p, c.c_int,
                                 _u32p, c.c_int]
        L.or_par_cc.restype = None
        L.or_par_cc.argtypes = [c.c_uint32, c.c_uint64, _u32p, _u32p, _u32p, c.c_void_p,
                                c.c_int]
        L.or_wkey.restype = c.c_uint32
        L.or_wkey.argtypes = [c.c_uint32, c.c_int]
        _lib = L
    return _lib


def _arr(x):
    return np.ascontiguousarray(x, dtype=np.uint32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def kruskal(n, u, v, w, wmode=1, labels=True):
    """Sequential Kruskal MSF: (sorted ids, labels or None)."""
    u, v, w = _arr(u), _arr(v), _arr(w)
    ids = np.empty(max(n, 1), np.uint32)
    lab = np.empty(n, np.uint32) if labels else None
    c = lib().or_kruskal(n, len(u), u, v, w, wmode, ids, _ptr(lab))
    if c < 0:
        raise MemoryError("or_kruskal failed")
    return ids[:c].copy(), lab


def boruvka(n, u, v, w, wmode=1, max_rounds=200):
    """Canonical Boruvka: dict(ids, labels, log_m, log_c, rounds)."""
    u, v, w = _arr(u), _arr(v), _arr(w)
    ids = np.empty(max(n, 1), np.uint32)
    lab = np.empty(n, np.uint32)
    lm = np.zeros(max_rounds, np.uint64)
    lc = np.zeros(max_rounds, np.uint64)
    r = ctypes.c_int(0)
    c = lib().or_boruvka(n, len(u), u, v, w, wmode, ids, _ptr(lab), _ptr(lm), _ptr(lc),
                         max_rounds, ctypes.byref(r))
    if c < 0:
        raise RuntimeError("or_boruvka: round cap hit or out of memory")
    return dict(ids=ids[:c].copy(), labels=lab, log_m=lm[:r.value].copy(),
                log_c=lc[:r.value].copy(), rounds=r.value)


def boruvka_rounds(n, u, v, w, wmode, rounds):
    """State of the canonical Boruvka loop after `rounds` rounds: (forest ids
    chosen so far, ascending; min-id labels of the components so far)."""
    u, v, w = _arr(u), _arr(v), _arr(w)
    ids = np.empty(max(n, 1), np.uint32)
    lab = np.empty(max(n, 1), np.uint32)
    c = lib().or_boruvka_rounds(n, len(u), u, v, w, wmode, rounds, ids, lab)
    if c < 0:
        raise MemoryError("or_boruvka_rounds failed")
    return ids[:c].copy(), lab[:n].copy()


def uf_labels(n, u, v, forest=False):
    u, v = _arr(u), _arr(v)
    lab = np.empty(n, np.uint32)
    f = np.empty(n, np.uint32) if forest else None
    lib().or_uf_labels(n, len(u), u, v, lab, _ptr(f))
    return (lab, f) if forest else lab


def bfs_labels(n, u, v):
    u, v = _arr(u), _arr(v)
    lab = np.empty(n, np.uint32)
    lib().or_bfs_labels(n, len(u), u, v, lab)
    return lab


def check_forest(n, u, v, ids, labels):
    u, v, ids, labels = _arr(u), _arr(v), _arr(ids), _arr(labels)
    return lib().or_check_forest(n, len(u), u, v, ids, len(ids), labels)


def weight_sum(ids, w, wmode=1):
    ids, w = _arr(ids), _arr(w)
    ex = ctypes.c_int(0)
    s = lib().or_weight_sum(len(ids), ids, w, wmode, ctypes.byref(ex))
    return int(s), bool(ex.value)


def par_mst(n, u, v, w, wmode=1, threads=0):
    u, v, w = _arr(u), _arr(v), _arr(w)
    ids = np.empty(max(n, 1), np.uint32)
    c = lib().or_par_mst(n, len(u), u, v, w, wmode, ids, threads)
    if c < 0:
        raise MemoryError("or_par_mst failed")
    return ids[:c].copy()


def par_cc(n, u, v, threads=0, forest=False):
    u, v = _arr(u), _arr(v)
    lab = np.empty(n, np.uint32)
    f = np.empty(n, np.uint32) if forest else None
    lib().or_par_cc(n, len(u), u, v, lab, _ptr(f), threads)
    return (lab, f) if forest else lab


def wkey(w, wmode=1):
    """Order key of weight bit patterns (or_wkey in phem_oracle.c): float32
    -0.0 equals +0.0, every NaN is one value above +inf."""
    w = np.asarray(w, dtype=np.uint32)
    if wmode == 0:
        return w.copy()
    neg = (w & np.uint32(0x80000000)) != 0
    k = np.where(neg, ~w, w | np.uint32(0x80000000)).astype(np.uint32)
    mag = w & np.uint32(0x7FFFFFFF)
    k = np.where(mag == 0, np.uint32(0x80000000), k)
    return np.where(mag > np.uint32(0x7F800000), np.uint32(0xFFFFFFFF), k).astype(np.uint32)


# ---- reference CPU builds (oracle/_ref), present only where they were built ----

def ref_available():
    return (os.path.exists(os.path.join(REF_DIR, "libref_serialmst.so"))
            and os.path.exists(os.path.join(REF_DIR, "libref_hostcc.so")))


_ref_mst = None
_ref_cc = None


def ref_serial_mst(n, u, v, w_f32):
    """Reference PBBS serial filter-Kruskal (float weights); ids sorted."""
    global _ref_mst
    if _ref_mst is None:
        L = ctypes.CDLL(os.path.join(REF_DIR, "libref_serialmst.so"))
        L.ref_serial_mst.restype = ctypes.c_long
        L.ref_serial_mst.argtypes = [ctypes.c_int, ctypes.c_long, _u32p, _u32p,
                                     np.ctypeslib.ndpointer(dtype=np.float32,
                                                            flags="C_CONTIGUOUS"),
                                     _u32p]
        _ref_mst = L
    u, v = _arr(u), _arr(v)
    w = np.ascontiguousarray(w_f32, dtype=np.float32)
    ids = np.empty(max(n, 1), np.uint32)
    c = _ref_mst.ref_serial_mst(n, len(u), u, v, w, ids)
    return np.sort(ids[:c])


def ref_host_cc(n, u, v):
    """Reference HAST CPU union-find over packed (u << 32 | v) edges:
    (labels, sptree) with sptree[root] = packed hooking edge or INT_MAX."""
    global _ref_cc
    if _ref_cc is None:
        L = ctypes.CDLL(os.path.join(REF_DIR, "libref_hostcc.so"))
        L.ref_host_cc.restype = ctypes.c_int
        L.ref_host_cc.argtypes = [ctypes.c_int, ctypes.c_long, _u64p, _u32p, _u64p]
        _ref_cc = L
    e = (np.asarray(u, np.uint64) << np.uint64(32)) | np.asarray(v, np.uint64)
    e = np.ascontiguousarray(e)
    lab = np.empty(n, np.uint32)
    sp = np.empty(n, np.uint64)
    _ref_cc.ref_host_cc(n, len(e), e, lab, sp)
    return lab, sp


def ref_host_cc_time(n, edges):
    """Seconds of the reference's HAST host_spanning_tree over the packed
    (u << 32 | v) records `edges` (numpy uint64), inputs prepared outside the
    timer; and the ParlayLib worker count it ran with."""
    global _ref_cc
    if _ref_cc is None:
        ref_host_cc(1, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    L = _ref_cc
    L.ref_host_cc_time.restype = ctypes.c_double
    L.ref_host_cc_time.argtypes = [ctypes.c_int, ctypes.c_long, _u64p]
    L.ref_host_cc_workers.restype = ctypes.c_int
    e = np.ascontiguousarray(edges, dtype=np.uint64)
    return L.ref_host_cc_time(n, len(e), e), L.ref_host_cc_workers()
