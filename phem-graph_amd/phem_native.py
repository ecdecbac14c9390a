"""Python host side of the PHEM MI355X library: ctypes bindings of libphem.so.

The graph work runs in the HIP kernels of libphem.so; torch only provides
device memory and the current HIP stream.  There is no CPU fallback: every
entry point raises PhemError when the library or a gfx950 device is missing.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphem.so")

W_U32 = 0
W_F32 = 1
NO_EDGE = 0xFFFFFFFFFFFFFFFF
MAX_ROUNDS = 64
# RMAT (a, b, c, d) = (0.57, 0.19, 0.19, 0.05) as cumulative 32-bit thresholds
RMAT_T = (2448131359, 3264175145, 4080218931)


class PhemError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


EINVAL, ENOMEM, EHIP, ENOCONV, ENODEV, EBUSY, ENOTFOUND = -1, -2, -3, -4, -5, -6, -7
REQ_DEVICE = 1
REQ_STREAM = 2


class MstOpts(ctypes.Structure):
    _fields_ = [("light_factor", ctypes.c_double), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class MstStats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_uint32),
        ("mst_count", ctypes.c_uint32),
        ("weight_fx", ctypes.c_uint64),
        ("weight_exact", ctypes.c_int32),
        ("light_rounds", ctypes.c_uint32),
        ("sum_m", ctypes.c_uint64),
        ("sum_c", ctypes.c_uint64),
        ("valid_edges", ctypes.c_uint64),
        ("tau", ctypes.c_uint64),
        ("light_edges", ctypes.c_uint64),
        ("heavy_edges", ctypes.c_uint64),
        ("giant_edges", ctypes.c_uint64),
        ("m_log", ctypes.c_uint64 * MAX_ROUNDS),
        ("c_log", ctypes.c_uint64 * MAX_ROUNDS),
    ]

    def as_dict(self):
        r = self.rounds
        return dict(rounds=r, mst_count=self.mst_count, weight_fx=self.weight_fx,
                    weight_exact=bool(self.weight_exact), sum_m=self.sum_m,
                    sum_c=self.sum_c, valid_edges=self.valid_edges, tau=self.tau,
                    light_edges=self.light_edges, heavy_edges=self.heavy_edges,
                    giant_edges=self.giant_edges,
                    light_rounds=self.light_rounds,
                    m_log=list(self.m_log[:r]),
                    c_log=list(self.c_log[:r]))


class ServerOpts(ctypes.Structure):
    _fields_ = [("mem_bytes", ctypes.c_uint64), ("max_batch_edges", ctypes.c_uint64),
                ("slots", ctypes.c_uint32), ("max_graphs", ctypes.c_uint32),
                ("window_us", ctypes.c_uint32), ("wmode", ctypes.c_int32),
                ("record_rounds", ctypes.c_uint32), ("stream_mem_mib", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("m", ctypes.c_uint64),
                ("u", ctypes.c_void_p), ("v", ctypes.c_void_p), ("w", ctypes.c_void_p),
                ("forest_ids", ctypes.c_void_p), ("labels", ctypes.c_void_p),
                ("arrival", ctypes.c_double), ("stream_id", ctypes.c_uint64)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("forest_count", ctypes.c_uint32),
                ("weight_fx", ctypes.c_uint64), ("weight_exact", ctypes.c_int32),
                ("batch_graphs", ctypes.c_uint32), ("arrival", ctypes.c_double),
                ("start", ctypes.c_double), ("done", ctypes.c_double),
                ("error", ctypes.c_char * 160)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["error"] = self.error.decode(errors="replace")
        return d


class ServerStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in
                ("submitted", "completed", "refused", "batches", "edges")] + \
               [(f, ctypes.c_double) for f in
                ("first_arrival", "last_done", "lat_p50", "lat_p99", "lat_max", "lat_mean",
                 "mean_batch_graphs")] + \
               [("slot_bytes", ctypes.c_uint64), ("slot_edges", ctypes.c_uint64),
                ("recorded_batches", ctypes.c_uint64), ("record_fallbacks", ctypes.c_uint64),
                ("recordings", ctypes.c_uint64), ("stream_batches", ctypes.c_uint64),
                ("stream_edges", ctypes.c_uint64), ("stream_lat_p50", ctypes.c_double),
                ("stream_lat_p99", ctypes.c_double), ("streams_open", ctypes.c_uint64),
                ("stream_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GraphFile(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("wmode", ctypes.c_int32), ("m", ctypes.c_uint64),
                ("u", ctypes.c_void_p), ("v", ctypes.c_void_p), ("w", ctypes.c_void_p),
                ("edges", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("mst_light_factor", ctypes.c_double), ("cc_sample_factor", ctypes.c_double),
                ("mst_block", ctypes.c_uint32), ("mst_ept", ctypes.c_uint32),
                ("mst_ept_compact", ctypes.c_uint32), ("mst_compact_log2", ctypes.c_uint32),
                ("mst_split_ratio", ctypes.c_double),
                ("cc_block", ctypes.c_uint32), ("cc_ept", ctypes.c_uint32),
                ("cc_ept_filter", ctypes.c_uint32), ("cc_block_filter", ctypes.c_uint32),
                ("cpu_share", ctypes.c_double), ("mst_light_factor_skewed", ctypes.c_double)]


class ChunksStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("cpu_seconds", ctypes.c_double),
                ("chunks", ctypes.c_uint64), ("cpu_edges", ctypes.c_uint64),
                ("cpu_kept", ctypes.c_uint64), ("reserved", ctypes.c_uint64 * 3)]


# the knobs of phem_tuning, in order (the reserved words excluded)
TUNING_KNOBS = ("blocks_per_cu", "mst_light_factor", "cc_sample_factor", "mst_block",
                "mst_ept", "mst_ept_compact", "mst_compact_log2", "mst_split_ratio",
                "cc_block", "cc_ept", "cc_ept_filter", "cpu_share", "mst_light_factor_skewed",
                "cc_block_filter")
_FLOAT_KNOBS = ("mst_light_factor", "cc_sample_factor", "mst_split_ratio", "cpu_share",
                "mst_light_factor_skewed")
CPU_EDGES_AUTO = 0xFFFFFFFFFFFFFFFF


_lib = None


def lib():
    """Load libphem.so (built in-tree by `make -C phem-graph_amd`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PhemError(f"libphem.so not built: {LIB_PATH}")
    L = ctypes.CDLL(LIB_PATH)
    c = ctypes
    vp, u32, u64, i32 = c.c_void_p, c.c_uint32, c.c_uint64, c.c_int
    sig = {
        "phem_version": (c.c_char_p, []),
        "phem_last_error": (c.c_char_p, []),
        "phem_device_check": (i32, [c.POINTER(c.c_int)]),
        "phem_mst_workspace_bytes": (i32, [u32, u64, c.POINTER(c.c_uint64)]),
        "phem_mst_solve": (i32, [u32, u64, vp, vp, vp, i32, vp, u64, vp, vp,
                                 c.POINTER(MstOpts), c.POINTER(MstStats), vp]),
        "phem_mst_solve_noalloc": (i32, [u32, u64, vp, vp, vp, i32, vp, vp,
                                         c.POINTER(MstOpts), c.POINTER(MstStats), vp]),
        "phem_mst_solve_aos": (i32, [u32, u64, vp, i32, vp, u64, vp, vp,
                                     c.POINTER(MstOpts), c.POINTER(MstStats), vp]),
        "phem_mst_graph_create": (i32, [u32, u64, vp, vp, vp, i32, vp, u64, vp, vp,
                                        c.POINTER(MstOpts), i32, vp, c.POINTER(vp)]),
        "phem_mst_graph_launch": (i32, [vp, vp]),
        "phem_mst_graph_stats": (i32, [vp, c.POINTER(MstStats), vp]),
        "phem_mst_graph_destroy": (i32, [vp]),
        "phem_cc_init": (i32, [u32, vp, vp, vp]),
        "phem_cc_union": (i32, [u32, vp, u64, vp, vp, vp]),
        "phem_cc_labels": (i32, [u32, vp, vp, vp]),
        "phem_cc_solve": (i32, [u32, vp, u64, vp, vp, vp, vp]),
        "phem_cc_merge_forests": (i32, [u32, vp, vp, vp, vp, vp, vp]),
        "phem_gen_uniform": (i32, [u32, u64, u64, u64, i32, vp, vp, vp, vp]),
        "phem_gen_rmat": (i32, [u32, u64, u64, u32, u32, u32, i32, vp, vp, vp, vp]),
        "phem_pack_edges": (i32, [vp, vp, u64, vp, vp]),
        "phem_gather_edges_aos": (i32, [vp, vp, u64, vp, vp]),
        "phem_stream_create": (i32, [u32, u64, i32, c.POINTER(vp)]),
        "phem_stream_destroy": (i32, [vp]),
        "phem_stream_apply": (i32, [vp, vp, vp, vp, u64, vp]),
        "phem_stream_forest": (i32, [vp, c.POINTER(c.c_uint64), vp, vp, vp, vp]),
        "phem_stream_copy_forest": (i32, [vp, vp, vp, vp, vp, vp]),
        "phem_stream_stats": (i32, [vp, c.POINTER(MstStats), c.POINTER(c.c_uint64),
                                    c.POINTER(c.c_uint64)]),
        "phem_stream_labels": (i32, [vp, vp, vp]),
        "phem_stream_cc_forest": (i32, [vp, c.POINTER(vp)]),
        "phem_stream_copy_cc_forest": (i32, [vp, vp, vp]),
        "phem_stream_snapshot_bytes": (i32, [vp, c.POINTER(c.c_uint64)]),
        "phem_stream_save": (i32, [vp, vp, u64, vp]),
        "phem_stream_restore": (i32, [vp, u64, u64, c.POINTER(vp)]),
        "phem_batch_workspace_bytes": (i32, [u64, u64, u32, c.POINTER(c.c_uint64)]),
        "phem_batch_solve": (i32, [u32, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64),
                                   vp, vp, vp, i32, vp, u64, vp, vp, vp, vp, vp,
                                   c.POINTER(MstOpts), c.POINTER(MstStats), vp]),
        "phem_batch_solve_ptrs": (i32, [u32, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64),
                                        c.POINTER(vp), c.POINTER(vp), c.POINTER(vp),
                                        i32, vp, u64, vp, vp, vp, vp, vp,
                                        c.POINTER(MstOpts), c.POINTER(MstStats), vp]),
        "phem_batch_graph_create": (i32, [u32, u64, u64, i32, vp, u64, vp, vp, vp, vp, vp,
                                          c.POINTER(MstOpts), i32, c.POINTER(vp)]),
        "phem_batch_graph_launch": (i32, [vp, u32, c.POINTER(c.c_uint64),
                                          c.POINTER(c.c_uint64), c.POINTER(vp),
                                          c.POINTER(vp), c.POINTER(vp), vp]),
        "phem_batch_graph_stats": (i32, [vp, c.POINTER(MstStats), vp]),
        "phem_batch_graph_destroy": (i32, [vp]),
        "phem_server_clock": (c.c_double, []),
        "phem_server_create": (i32, [c.POINTER(ServerOpts), c.POINTER(vp)]),
        "phem_server_destroy": (i32, [vp]),
        "phem_server_submit": (i32, [vp, c.POINTER(Request), c.POINTER(c.c_uint64)]),
        "phem_server_wait": (i32, [vp, u64, c.POINTER(Result)]),
        "phem_server_poll": (i32, [vp, u64, c.POINTER(Result), c.POINTER(c.c_int)]),
        "phem_server_drain": (i32, [vp]),
        "phem_server_get_stats": (i32, [vp, c.POINTER(ServerStats)]),
        "phem_server_reset_stats": (i32, [vp]),
        "phem_server_stream_open": (i32, [vp, u32, u64, c.POINTER(c.c_uint64)]),
        "phem_server_stream_close": (i32, [vp, u64]),
        "phem_stream_bytes": (i32, [u32, u64, c.POINTER(c.c_uint64)]),
        "phem_get_tuning": (i32, [c.POINTER(Tuning)]),
        "phem_set_tuning": (i32, [c.POINTER(Tuning)]),
        "phem_cpu_msf": (i32, [u32, u64, vp, vp, vp, i32, vp, c.POINTER(c.c_uint64), i32]),
        "phem_msf_host_chunks": (i32, [u32, u64, vp, vp, vp, i32, u64, u64, vp,
                                       c.POINTER(c.c_uint64), vp]),
        "phem_cpu_cc": (i32, [u32, u64, vp, vp, vp, i32]),
        "phem_read_mtx": (i32, [c.c_char_p, i32, c.POINTER(GraphFile)]),
        "phem_read_edge_list": (i32, [c.c_char_p, c.POINTER(GraphFile)]),
        "phem_read_ecl": (i32, [c.c_char_p, c.POINTER(GraphFile)]),
        "phem_read_pbbs_weighted": (i32, [c.c_char_p, c.POINTER(GraphFile)]),
        "phem_write_mtx_weighted": (i32, [c.c_char_p, c.POINTER(GraphFile)]),
        "phem_graph_file_free": (None, [c.POINTER(GraphFile)]),
        "phem_prof_enable": (i32, [i32]),
        "phem_prof_only": (i32, [c.c_char_p]),
        "phem_prof_reset": (i32, []),
        "phem_prof_count": (i32, []),
        "phem_prof_get": (i32, [i32, c.c_char_p, i32, c.POINTER(c.c_uint64),
                                c.POINTER(c.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = lib().phem_last_error().decode(errors="replace")
        raise PhemError(f"{what} returned {rc}: {msg}", rc)


def require_device():
    """Raise unless a gfx950 device is visible to the library."""
    cnt = ctypes.c_int(0)
    _check(lib().phem_device_check(ctypes.byref(cnt)), "phem_device_check")


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def as_u32(t):
    """Reinterpret a 32-bit tensor as uint32 (no copy)."""
    return t.view(torch.uint32) if t.dtype != torch.uint32 else t


def mst_workspace_bytes(n, m):
    b = ctypes.c_uint64(0)
    _check(lib().phem_mst_workspace_bytes(n, m, ctypes.byref(b)), "phem_mst_workspace_bytes")
    return b.value


MST_MAXIMUM = 1


def _opts(light_factor, maximum):
    if light_factor is None and not maximum:
        return None
    lf = float("nan") if light_factor is None else light_factor
    return ctypes.byref(MstOpts(lf, MST_MAXIMUM if maximum else 0, 0))


class MstSolver:
    """Reusable workspace + outputs for phem_mst_solve on one device.

    Inputs are uint32 / int32 device tensors u, v, w (w holds float32 or
    uint32 bit patterns).  Returns (ids, labels, stats): the forest's edge ids
    ascending (int32 view of uint32), min-id labels, and the round log."""

    def __init__(self, n, m, device="cuda"):
        self.n, self.m, self.device = n, m, torch.device(device)
        self.ws_bytes = mst_workspace_bytes(n, m)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)
        self.ids = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)
        self.labels = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)

    def solve(self, u, v, w, wmode=W_F32, labels=True, stream=None, light_factor=None,
              n=None, maximum=False):
        """light_factor: None = library default; <= 0 = canonical Boruvka.
        n: vertex count of this graph (default and maximum: the solver's).
        maximum: maximum spanning forest ((weight, id) descending)."""
        m = u.numel()
        n = self.n if n is None else n
        if m > self.m or n > self.n:
            raise PhemError(f"n={n} m={m} exceed solver capacity {self.n}, {self.m}")
        st = MstStats()
        opts = _opts(light_factor, maximum)
        _check(lib().phem_mst_solve(n, m, _p(u), _p(v), _p(w), wmode, _p(self.ws),
                                    self.ws_bytes, _p(self.ids),
                                    _p(self.labels) if labels else None, opts,
                                    ctypes.byref(st), _stream(stream)), "phem_mst_solve")
        s = st.as_dict()
        return self.ids[: s["mst_count"]], (self.labels[:n] if labels else None), s

    def solve_aos(self, edges, wmode=W_F32, labels=True, stream=None, light_factor=None,
                  maximum=False):
        """edges: device tensor of m x 3 32-bit words {u, v, w}."""
        m = edges.numel() // 3
        if m > self.m:
            raise PhemError(f"m={m} exceeds solver capacity {self.m}")
        st = MstStats()
        opts = _opts(light_factor, maximum)
        _check(lib().phem_mst_solve_aos(self.n, m, _p(edges), wmode, _p(self.ws),
                                        self.ws_bytes, _p(self.ids),
                                        _p(self.labels) if labels else None, opts,
                                        ctypes.byref(st), _stream(stream)),
               "phem_mst_solve_aos")
        s = st.as_dict()
        return self.ids[: s["mst_count"]], (self.labels[: self.n] if labels else None), s


class MstGraph:
    """phem_mst_graph_*: one solve of the device buffers u, v, w recorded into
    a HIP graph, replayed by launch(); stats() waits for the replay and
    returns the round log (PhemError with PHEM_ENOCONV when a phase needed
    more than its recorded rounds).  The buffers' contents may change between
    replays; their addresses and sizes may not."""

    def __init__(self, n, u, v, w, wmode=W_F32, rounds=0, light_factor=None,
                 maximum=False, labels=True, stream=None):
        self.n, self.m = n, u.numel()
        dev = u.device
        self.inputs = (u, v, w)  # kept alive: the graph holds their addresses
        self.ws_bytes = mst_workspace_bytes(n, self.m)
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        self.ids = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        self.labels = torch.empty(max(n, 1), dtype=torch.int32, device=dev) if labels else None
        h = ctypes.c_void_p()
        _check(lib().phem_mst_graph_create(n, self.m, _p(u), _p(v), _p(w), wmode,
                                           _p(self.ws), self.ws_bytes, _p(self.ids),
                                           _p(self.labels), _opts(light_factor, maximum),
                                           rounds, _stream(stream), ctypes.byref(h)),
               "phem_mst_graph_create")
        self.h = h

    def launch(self, stream=None):
        _check(lib().phem_mst_graph_launch(self.h, _stream(stream)), "phem_mst_graph_launch")

    def stats(self, stream=None):
        st = MstStats()
        _check(lib().phem_mst_graph_stats(self.h, ctypes.byref(st), _stream(stream)),
               "phem_mst_graph_stats")
        return st.as_dict()

    def result(self, stream=None):
        """(ids, labels, stats) of the last replay."""
        s = self.stats(stream)
        lab = self.labels[: self.n] if self.labels is not None else None
        return self.ids[: s["mst_count"]], lab, s

    def close(self):
        if self.h:
            lib().phem_mst_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mst_solve_noalloc(n, u, v, w, ids, labels=None, wmode=W_F32, light_factor=None,
                      maximum=False, stream=None):
    """phem_mst_solve_noalloc into the caller's device tensors ids (n slots)
    and labels (n slots or None); returns the stats dict."""
    st = MstStats()
    _check(lib().phem_mst_solve_noalloc(n, u.numel(), _p(u), _p(v), _p(w), wmode, _p(ids),
                                        _p(labels), _opts(light_factor, maximum),
                                        ctypes.byref(st), _stream(stream)),
           "phem_mst_solve_noalloc")
    return st.as_dict()


def mst_solve(n, u, v, w, wmode=W_F32, labels=True, light_factor=None, maximum=False):
    return MstSolver(n, u.numel(), device=u.device).solve(u, v, w, wmode, labels,
                                                          light_factor=light_factor,
                                                          maximum=maximum)


def cc_solve(n, edges, forest=True, stream=None):
    """Union-find CC over packed (u << 32 | v) int64 edges.
    Returns (labels, parent, forest or None)."""
    dev = edges.device
    par = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    lab = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    f = torch.empty(max(n, 1), dtype=torch.int64, device=dev) if forest else None
    _check(lib().phem_cc_solve(n, _p(edges), edges.numel(), _p(par), _p(f), _p(lab),
                               _stream(stream)), "phem_cc_solve")
    return lab[:n], par[:n], (f[:n] if forest else None)


def cc_init(n, parent, forest=None, stream=None):
    _check(lib().phem_cc_init(n, _p(parent), _p(forest), _stream(stream)), "phem_cc_init")


def cc_union(n, edges, parent, forest=None, stream=None):
    _check(lib().phem_cc_union(n, _p(edges), edges.numel(), _p(parent), _p(forest),
                               _stream(stream)), "phem_cc_union")


def cc_labels(n, parent, labels, stream=None):
    _check(lib().phem_cc_labels(n, _p(parent), _p(labels), _stream(stream)), "phem_cc_labels")


def cc_merge_forests(n, forest, extra, parent, forest2, labels, stream=None):
    _check(lib().phem_cc_merge_forests(n, _p(forest), _p(extra), _p(parent), _p(forest2),
                                       _p(labels), _stream(stream)), "phem_cc_merge_forests")


def gen_uniform(n, m, seed, positive=False, id0=0, device="cuda", out=None):
    """m uniform edges on n vertices: int32 device tensors (u, v, w-bits)."""
    if out is None:
        out = tuple(torch.empty(m, dtype=torch.int32, device=device) for _ in range(3))
    u, v, w = out
    _check(lib().phem_gen_uniform(n, m, seed, id0, int(positive), _p(u), _p(v), _p(w),
                                  _stream()), "phem_gen_uniform")
    return u, v, w


def gen_rmat(scale, m, seed, t=RMAT_T, positive=False, device="cuda", out=None):
    if out is None:
        out = tuple(torch.empty(m, dtype=torch.int32, device=device) for _ in range(3))
    u, v, w = out
    _check(lib().phem_gen_rmat(scale, m, seed, t[0], t[1], t[2], int(positive), _p(u),
                               _p(v), _p(w), _stream()), "phem_gen_rmat")
    return u, v, w


def pack_edges(u, v, out=None):
    e = out if out is not None else torch.empty(u.numel(), dtype=torch.int64, device=u.device)
    _check(lib().phem_pack_edges(_p(u), _p(v), u.numel(), _p(e), _stream()), "phem_pack_edges")
    return e


def prof_enable(on=True):
    lib().phem_prof_enable(int(on))


def prof_only(name=None):
    """Record only kernel `name` (None: every kernel)."""
    lib().phem_prof_only(name.encode() if name else None)


def prof_reset():
    lib().phem_prof_reset()


def prof_read():
    """{kernel name: (launches, total ms)} since the last reset (waits for the
    recorded events)."""
    L = lib()
    out = {}
    for i in range(L.phem_prof_count()):
        name = ctypes.create_string_buffer(64)
        n = ctypes.c_uint64(0)
        ms = ctypes.c_double(0)
        _check(L.phem_prof_get(i, name, 64, ctypes.byref(n), ctypes.byref(ms)), "phem_prof_get")
        out[name.value.decode()] = (n.value, ms.value)
    return out


def batch_solve(voff, eoff, u, v, w, wmode=W_F32, labels=True, weights=True,
                stream=None, light_factor=None, ws=None, maximum=False):
    """k independent graphs in one solve (phem_batch_solve).

    voff / eoff: k + 1 offsets (python ints) into the concatenated vertex and
    edge ranges; u, v, w: concatenated int32 device tensors with graph-local
    vertex ids.  Returns a dict with per-graph lists: ids (local forest edge
    ids, ascending), labels (local min-id labels), weight_fx and inexact, plus
    the solve's stats."""
    k = len(voff) - 1
    n_total, m_total = int(voff[-1]), int(eoff[-1])
    dev = u.device
    hv = (ctypes.c_uint64 * (k + 1))(*[int(x) for x in voff])
    he = (ctypes.c_uint64 * (k + 1))(*[int(x) for x in eoff])
    need = ctypes.c_uint64(0)
    _check(lib().phem_batch_workspace_bytes(n_total, m_total, k, ctypes.byref(need)),
           "phem_batch_workspace_bytes")
    if ws is None or ws.numel() < need.value:
        ws = torch.empty(max(need.value, 1), dtype=torch.uint8, device=dev)
    ids = torch.empty(max(n_total, 1), dtype=torch.int32, device=dev)
    cnt = torch.empty(k, dtype=torch.int32, device=dev)
    lab = torch.empty(max(n_total, 1), dtype=torch.int32, device=dev) if labels else None
    wfx = torch.empty(k, dtype=torch.int64, device=dev) if weights else None
    bad = torch.empty(k, dtype=torch.int64, device=dev) if weights else None
    opts = _opts(light_factor, maximum)
    st = MstStats()
    _check(lib().phem_batch_solve(k, hv, he, _p(u), _p(v), _p(w), wmode, _p(ws),
                                  ws.numel(), _p(ids), _p(cnt), _p(lab), _p(wfx), _p(bad),
                                  opts, ctypes.byref(st), _stream(stream)),
           "phem_batch_solve")
    cnt_h = cnt.cpu().tolist()
    ids_h = ids.cpu()
    lab_h = lab.cpu() if labels else None
    out = {"ids": [], "labels": [], "stats": st.as_dict(), "counts": cnt_h}
    for g in range(k):
        a = int(voff[g])
        out["ids"].append(ids_h[a: a + cnt_h[g]])
        if labels:
            out["labels"].append(lab_h[a: int(voff[g + 1])])
    if weights:
        out["weight_fx"] = [x & 0xFFFFFFFFFFFFFFFF for x in wfx.cpu().tolist()]
        out["inexact"] = bad.cpu().tolist()
    return out


class BatchGraph:
    """phem_batch_graph_*: a batched solve recorded for up to k_cap - 1 graphs
    of n_cap vertices and m_cap edges in total, replayed by launch() for any
    batch that fits; results() waits for the replay and returns the same
    dict as batch_solve (PhemError with PHEM_ENOCONV when a phase needed more
    than the recorded rounds)."""

    def __init__(self, k_cap, n_cap, m_cap, wmode=W_F32, rounds=0, light_factor=None,
                 labels=True, weights=True, device="cuda"):
        dev = torch.device(device)
        need = ctypes.c_uint64(0)
        _check(lib().phem_batch_workspace_bytes(n_cap, m_cap, k_cap, ctypes.byref(need)),
               "phem_batch_workspace_bytes")
        self.k_cap, self.n_cap, self.m_cap = k_cap, n_cap, m_cap
        self.ws = torch.empty(max(need.value, 1), dtype=torch.uint8, device=dev)
        self.ids = torch.empty(max(n_cap, 1), dtype=torch.int32, device=dev)
        self.cnt = torch.empty(k_cap, dtype=torch.int32, device=dev)
        self.lab = torch.empty(max(n_cap, 1), dtype=torch.int32, device=dev) if labels else None
        self.wfx = torch.empty(k_cap, dtype=torch.int64, device=dev) if weights else None
        self.bad = torch.empty(k_cap, dtype=torch.int64, device=dev) if weights else None
        h = ctypes.c_void_p()
        _check(lib().phem_batch_graph_create(k_cap, n_cap, m_cap, wmode, _p(self.ws),
                                             self.ws.numel(), _p(self.ids), _p(self.cnt),
                                             _p(self.lab), _p(self.wfx), _p(self.bad),
                                             _opts(light_factor, False), rounds,
                                             ctypes.byref(h)), "phem_batch_graph_create")
        self.h = h
        self.voff = None
        self.keep = None

    def launch(self, graphs, stream=None):
        """graphs: list of (n, u, v, w) with int32 device tensors u, v, w."""
        k = len(graphs)
        voff, eoff = [0], [0]
        for n, u, _, _ in graphs:
            voff.append(voff[-1] + n)
            eoff.append(eoff[-1] + u.numel())
        tab = [(ctypes.c_void_p * max(k, 1))(*[g[i].data_ptr() if g[1].numel() else None
                                              for g in graphs]) for i in (1, 2, 3)]
        hv = (ctypes.c_uint64 * (k + 1))(*voff)
        he = (ctypes.c_uint64 * (k + 1))(*eoff)
        _check(lib().phem_batch_graph_launch(self.h, k, hv, he, tab[0], tab[1], tab[2],
                                             _stream(stream)), "phem_batch_graph_launch")
        self.voff, self.keep = voff, graphs

    def stats(self, stream=None):
        st = MstStats()
        _check(lib().phem_batch_graph_stats(self.h, ctypes.byref(st), _stream(stream)),
               "phem_batch_graph_stats")
        return st.as_dict()

    def results(self, stream=None):
        st = self.stats(stream)
        k = len(self.voff) - 1
        cnt = self.cnt[:k].cpu().tolist()
        ids_h, lab_h = self.ids.cpu(), (self.lab.cpu() if self.lab is not None else None)
        out = {"ids": [], "labels": [], "stats": st, "counts": cnt}
        for g in range(k):
            a = self.voff[g]
            out["ids"].append(ids_h[a: a + cnt[g]])
            if lab_h is not None:
                out["labels"].append(lab_h[a: self.voff[g + 1]])
        if self.wfx is not None:
            out["weight_fx"] = [x & 0xFFFFFFFFFFFFFFFF for x in self.wfx[:k].cpu().tolist()]
            out["inexact"] = self.bad[:k].cpu().tolist()
        return out

    def close(self):
        if self.h is not None and self.h.value:
            lib().phem_batch_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    """Streaming minimum spanning forest + connected components with the
    forest and parent array resident in HBM (phem_stream_*)."""

    def __init__(self, n, max_batch, wmode=W_F32, device="cuda"):
        self.n, self.device = n, torch.device(device)
        h = ctypes.c_void_p()
        _check(lib().phem_stream_create(n, max_batch, wmode, ctypes.byref(h)),
               "phem_stream_create")
        self.h = h

    def close(self):
        if self.h is not None and self.h.value:
            lib().phem_stream_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, u, v, w, stream=None):
        _check(lib().phem_stream_apply(self.h, _p(u), _p(v), _p(w), u.numel(),
                                       _stream(stream)), "phem_stream_apply")

    def stats(self):
        st = MstStats()
        e = ctypes.c_uint64(0)
        b = ctypes.c_uint64(0)
        _check(lib().phem_stream_stats(self.h, ctypes.byref(st), ctypes.byref(e),
                                       ctypes.byref(b)), "phem_stream_stats")
        d = st.as_dict()
        d["edges_inserted"], d["batches"] = e.value, b.value
        return d

    def forest(self):
        """(u, v, w, gid) device tensors of the current forest (copies)."""
        cnt = ctypes.c_uint64(0)
        _check(lib().phem_stream_forest(self.h, ctypes.byref(cnt), None, None, None, None),
               "phem_stream_forest")
        k = cnt.value
        out = [torch.empty(max(k, 1), dtype=torch.int32, device=self.device) for _ in range(3)]
        g = torch.empty(max(k, 1), dtype=torch.int64, device=self.device)
        _check(lib().phem_stream_copy_forest(self.h, _p(out[0]), _p(out[1]), _p(out[2]),
                                             _p(g), _stream()), "phem_stream_copy_forest")
        return out[0][:k], out[1][:k], out[2][:k], g[:k]

    def labels(self):
        lab = torch.empty(max(self.n, 1), dtype=torch.int32, device=self.device)
        _check(lib().phem_stream_labels(self.h, _p(lab), _stream()), "phem_stream_labels")
        return lab[: self.n]

    def save(self, stream=None):
        """Host snapshot (numpy uint8 array) of the stream's state."""
        import numpy as np
        b = ctypes.c_uint64(0)
        _check(lib().phem_stream_snapshot_bytes(self.h, ctypes.byref(b)),
               "phem_stream_snapshot_bytes")
        buf = np.zeros(b.value, np.uint8)
        _check(lib().phem_stream_save(self.h, buf.ctypes.data, b.value, _stream(stream)),
               "phem_stream_save")
        return buf

    @classmethod
    def restore(cls, snapshot, max_batch, device="cuda"):
        """A stream resumed from a save() snapshot (numpy uint8 array)."""
        import numpy as np
        snap = np.ascontiguousarray(snapshot, np.uint8)
        h = ctypes.c_void_p()
        _check(lib().phem_stream_restore(snap.ctypes.data if snap.size else None,
                                         snap.size, max_batch, ctypes.byref(h)),
               "phem_stream_restore")
        self = cls.__new__(cls)
        self.device = torch.device(device)
        self.h = h
        n = ctypes.c_uint32.from_buffer_copy(snap[8:12].tobytes()).value
        self.n = n
        return self

    def cc_forest(self):
        """Copy of the union-find forest: n packed (u << 32 | v) slots, all
        ones where no edge hooked that vertex."""
        f = torch.empty(max(self.n, 1), dtype=torch.int64, device=self.device)
        _check(lib().phem_stream_copy_cc_forest(self.h, _p(f), _stream()),
               "phem_stream_copy_cc_forest")
        return f[: self.n]


def _host_ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None and a.size else None


class Server:
    """Request loop with continuous batching (phem_server_*).

    submit() takes a graph either as int32 device tensors (u, v, w) or as
    numpy uint32 host arrays, plus optional numpy uint32 host arrays that
    receive the forest edge ids (n slots) and labels (n slots).  The arrays
    are kept referenced until the request is waited for."""

    def __init__(self, mem_bytes=0, slots=2, max_graphs=4096, max_batch_edges=0,
                 window_us=0, wmode=W_F32, record_rounds=0, stream_mem_mib=0):
        o = ServerOpts(mem_bytes, max_batch_edges, slots, max_graphs, window_us, wmode,
                       record_rounds, stream_mem_mib)
        h = ctypes.c_void_p()
        _check(lib().phem_server_create(ctypes.byref(o), ctypes.byref(h)), "phem_server_create")
        self.h = h
        self.keep = {}

    @staticmethod
    def clock():
        return lib().phem_server_clock()

    def submit(self, n, u, v, w, forest_ids=None, labels=None, arrival=-1.0):
        dev = isinstance(u, torch.Tensor)
        if dev:
            ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t.numel() else None
            m = u.numel()
        else:
            ptr = _host_ptr
            m = u.size
        r = Request(n, REQ_DEVICE if dev else 0, m, ptr(u), ptr(v), ptr(w),
                    _host_ptr(forest_ids), _host_ptr(labels), arrival, 0)
        t = ctypes.c_uint64(0)
        _check(lib().phem_server_submit(self.h, ctypes.byref(r), ctypes.byref(t)),
               "phem_server_submit")
        self.keep[t.value] = (u, v, w, forest_ids, labels)
        return t.value

    def stream_open(self, n, max_batch):
        """A resident stream of n vertices (batches of up to max_batch edges);
        returns its id."""
        sid = ctypes.c_uint64(0)
        _check(lib().phem_server_stream_open(self.h, n, max_batch, ctypes.byref(sid)),
               "phem_server_stream_open")
        return sid.value

    def stream_close(self, sid):
        _check(lib().phem_server_stream_close(self.h, sid), "phem_server_stream_close")

    def submit_batch(self, sid, u, v, w, labels=None, arrival=-1.0):
        """A batch of edges (int32 device tensors or numpy uint32 host arrays)
        for resident stream sid; labels: optional numpy uint32 array of the
        stream's n slots receiving the labels after the batch."""
        dev = isinstance(u, torch.Tensor)
        if dev:
            ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t.numel() else None
            m = u.numel()
        else:
            ptr = _host_ptr
            m = u.size
        r = Request(0, (REQ_DEVICE if dev else 0) | REQ_STREAM, m, ptr(u), ptr(v), ptr(w),
                    None, _host_ptr(labels), arrival, sid)
        t = ctypes.c_uint64(0)
        _check(lib().phem_server_submit(self.h, ctypes.byref(r), ctypes.byref(t)),
               "phem_server_submit")
        self.keep[t.value] = (u, v, w, labels)
        return t.value

    def wait(self, ticket):
        res = Result()
        _check(lib().phem_server_wait(self.h, ticket, ctypes.byref(res)), "phem_server_wait")
        self.keep.pop(ticket, None)
        return res.as_dict()

    def poll(self, ticket):
        res = Result()
        done = ctypes.c_int(0)
        _check(lib().phem_server_poll(self.h, ticket, ctypes.byref(res), ctypes.byref(done)),
               "phem_server_poll")
        if done.value:
            self.keep.pop(ticket, None)
            return res.as_dict()
        return None

    def drain(self):
        _check(lib().phem_server_drain(self.h), "phem_server_drain")

    def stats(self):
        st = ServerStats()
        _check(lib().phem_server_get_stats(self.h, ctypes.byref(st)), "phem_server_get_stats")
        return st.as_dict()

    def reset_stats(self):
        _check(lib().phem_server_reset_stats(self.h), "phem_server_reset_stats")

    def close(self):
        if self.h is not None and self.h.value:
            lib().phem_server_destroy(self.h)
        self.h = None
        self.keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _take_graph(g):
    """Copies a GraphFile's arrays into numpy and frees the library's."""
    import numpy as np

    def arr(ptr, dt, k):
        if not ptr or k == 0:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(
            ctypes.c_uint32 if dt == np.uint32 else ctypes.c_uint64)), shape=(k,)).copy()
    out = {"n": g.n, "m": g.m, "wmode": g.wmode}
    if g.u:
        out["u"], out["v"] = arr(g.u, np.uint32, g.m), arr(g.v, np.uint32, g.m)
        out["w"] = arr(g.w, np.uint32, g.m)
    if g.edges:
        out["edges"] = arr(g.edges, np.uint64, g.m)
    lib().phem_graph_file_free(ctypes.byref(g))
    return out


def read_mtx(path, generate_weights=False):
    g = GraphFile()
    _check(lib().phem_read_mtx(os.fsencode(path), int(generate_weights), ctypes.byref(g)),
           "phem_read_mtx")
    return _take_graph(g)


def read_edge_list(path):
    g = GraphFile()
    _check(lib().phem_read_edge_list(os.fsencode(path), ctypes.byref(g)), "phem_read_edge_list")
    return _take_graph(g)


def read_ecl(path):
    g = GraphFile()
    _check(lib().phem_read_ecl(os.fsencode(path), ctypes.byref(g)), "phem_read_ecl")
    return _take_graph(g)


def read_pbbs_weighted(path):
    g = GraphFile()
    _check(lib().phem_read_pbbs_weighted(os.fsencode(path), ctypes.byref(g)),
           "phem_read_pbbs_weighted")
    return _take_graph(g)


def write_mtx_weighted(path, n, u, v, w):
    import numpy as np
    u, v, w = (np.ascontiguousarray(x, dtype=np.uint32) for x in (u, v, w))
    g = GraphFile(n, W_U32, len(u), u.ctypes.data, v.ctypes.data, w.ctypes.data, None)
    _check(lib().phem_write_mtx_weighted(os.fsencode(path), ctypes.byref(g)),
           "phem_write_mtx_weighted")


def cpu_msf(n, u, v, w, wmode=W_F32, threads=0):
    """Host minimum spanning forest (phem_cpu_msf): ascending edge ids."""
    import numpy as np
    u, v, w = (np.ascontiguousarray(x, dtype=np.uint32) for x in (u, v, w))
    ids = np.zeros(max(n, 1), np.uint32)
    cnt = ctypes.c_uint64(0)
    _check(lib().phem_cpu_msf(n, len(u), u.ctypes.data, v.ctypes.data, w.ctypes.data, wmode,
                              ids.ctypes.data, ctypes.byref(cnt), threads), "phem_cpu_msf")
    return ids[:cnt.value].copy()


def msf_host_chunks(n, u, v, w, wmode=W_F32, chunk_edges=1 << 24, cpu_edges=0):
    """Minimum spanning forest of host edges streamed through the device in
    chunks (phem_msf_host_chunks): (ascending edge ids, stats dict).
    cpu_edges: edges of the last chunk the host reduces meanwhile
    (CPU_EDGES_AUTO: the cpu_share knob's rule)."""
    import numpy as np
    u, v, w = (np.ascontiguousarray(x, dtype=np.uint32) for x in (u, v, w))
    ids = np.zeros(max(n, 1), np.uint32)
    cnt = ctypes.c_uint64(0)
    st = ChunksStats()
    _check(lib().phem_msf_host_chunks(n, len(u), u.ctypes.data, v.ctypes.data, w.ctypes.data,
                                      wmode, chunk_edges, cpu_edges, ids.ctypes.data,
                                      ctypes.byref(cnt), ctypes.byref(st)),
           "phem_msf_host_chunks")
    d = {f: getattr(st, f) for f, _ in ChunksStats._fields_ if f != "reserved"}
    return ids[:cnt.value].copy(), d


def get_tuning():
    t = Tuning()
    _check(lib().phem_get_tuning(ctypes.byref(t)), "phem_get_tuning")
    return {k: getattr(t, k) for k in TUNING_KNOBS}


def set_tuning(**knobs):
    """Sets the process-wide knobs named in TUNING_KNOBS (a knob left out, or
    given as None / 0, is unchanged)."""
    bad = set(knobs) - set(TUNING_KNOBS)
    if bad:
        raise PhemError(f"unknown tuning knobs {sorted(bad)}", EINVAL)
    t = Tuning()
    for k in TUNING_KNOBS:
        x = knobs.get(k)
        if k in _FLOAT_KNOBS:
            setattr(t, k, float("nan") if x is None else float(x))
        else:
            setattr(t, k, int(x or 0))
    _check(lib().phem_set_tuning(ctypes.byref(t)), "phem_set_tuning")


def cpu_cc(n, u, v, threads=0):
    """Host connected components (phem_cpu_cc): min-id labels."""
    import numpy as np
    u, v = (np.ascontiguousarray(x, dtype=np.uint32) for x in (u, v))
    lab = np.zeros(max(n, 1), np.uint32)
    _check(lib().phem_cpu_cc(n, len(u), u.ctypes.data, v.ctypes.data, lab.ctypes.data,
                             threads), "phem_cpu_cc")
    return lab[:n]
