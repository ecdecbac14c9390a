"""GPU tests of the C++ drop-in entry points (reference names and argument
meaning) in libphem.so:

- boruvka_for_edge_array / device_boruvka (through tests/harness) against the
  oracle's Kruskal forest, as the reference's `edge {u, v, w}` records;
- the reference's CC driver flow (its headers and its ParlayLib CPU path,
  CC/host_spanning_tree.cu, built in oracle/_ref) calling this library's
  SpanningTree and gpu_spanning_tree, for several GPU shares and batch sizes.
"""
import ctypes
import os

import numpy as np
import pytest

import cases
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "harness", "libdropin_mst.so")
HAST = os.path.join(ROOT, "oracle", "_ref", "libref_hast.so")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def H():
    import phem_native
    phem_native.require_device()
    if not os.path.exists(HARNESS):
        pytest.fail(f"drop-in harness not built: {HARNESS}")
    L = ctypes.CDLL(HARNESS)
    L.harness_boruvka_for_edge_array.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32,
                                                 _u32p, _u32p]
    L.harness_device_boruvka.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                         _u32p, _u32p, _u32p]
    return L


MST_CASES = ["u32_weights", "rmat_small", "grid_small", "duplicates_equal_w",
             "path_chords_small", "many_components", "out_of_range", "empty"]


def expected_records(n, u, v, w, wmode):
    ids, lab = O.kruskal(n, u, v, w, wmode)
    rec = np.stack([u[ids], v[ids], w[ids]], axis=1).astype(np.uint32)
    return ids, rec, lab


@pytest.mark.parametrize("name", MST_CASES)
def test_boruvka_for_edge_array(H, name):
    n, u, v, w, wmode = cases.make(name)
    # the reference's weights are unsigned integers; non-negative float bit
    # patterns order the same way as unsigned integers
    if wmode == 1 and (w & np.uint32(0x80000000)).any():
        pytest.skip("negative floats are not unsigned weights")
    aos = np.ascontiguousarray(np.stack([u, v, w], axis=1).reshape(-1), np.uint32)
    out = np.zeros(3 * max(n, 1), np.uint32)
    filled = np.array([0], np.uint32)
    assert H.harness_boruvka_for_edge_array(aos, n, len(u), out, filled) == 0
    ids, rec, _ = expected_records(n, u, v, w, 0)
    assert filled[0] == len(ids)
    assert np.array_equal(out[: 3 * len(ids)].reshape(-1, 3), rec)


@pytest.mark.parametrize("name", MST_CASES)
def test_device_boruvka_full_signature(H, name):
    n, u, v, w, wmode = cases.make(name)
    if wmode == 1 and (w & np.uint32(0x80000000)).any():
        pytest.skip("negative floats are not unsigned weights")
    aos = np.ascontiguousarray(np.stack([u, v, w], axis=1).reshape(-1), np.uint32)
    out = np.zeros(3 * max(n, 1), np.uint32)
    labels = np.zeros(max(n, 1), np.uint32)
    filled = np.array([0], np.uint32)
    m_after = np.array([7], np.uint32)
    assert H.harness_device_boruvka(aos, n, len(u), out, labels, filled, m_after) == 0
    ids, rec, lab = expected_records(n, u, v, w, 0)
    assert filled[0] == len(ids) and m_after[0] == 0
    assert np.array_equal(out[: 3 * len(ids)].reshape(-1, 3), rec)
    assert np.array_equal(labels[:n], lab)


@pytest.fixture(scope="module")
def HAST_LIB():
    import phem_native
    phem_native.require_device()
    if not os.path.exists(HAST):
        pytest.skip("oracle/_ref/libref_hast.so not built (needs the reference tree)")
    L = ctypes.CDLL(HAST)
    L.ref_hast_run.restype = ctypes.c_int
    L.ref_hast_run.argtypes = [ctypes.c_int, ctypes.c_long, _u64p, ctypes.c_float,
                               ctypes.c_int, _u32p, _u64p, _u64p,
                               ctypes.POINTER(ctypes.c_float)]
    return L


@pytest.mark.parametrize("share,batch", [(1.0, 1 << 20), (1.0, 4096), (0.6, 8192),
                                         (0.25, 1 << 20)])
@pytest.mark.parametrize("name", ["uniform_c1", "rmat_small", "grid_small", "many_components"])
def test_reference_cc_driver_over_dropin(HAST_LIB, name, share, batch):
    n, u, v, w, wmode = cases.make(name)
    rec = ((np.minimum(u, v).astype(np.uint64) << np.uint64(32))
           | np.maximum(u, v).astype(np.uint64))
    rec = np.ascontiguousarray(rec)
    labels = np.zeros(n, np.uint32)
    t1 = np.zeros(n, np.uint64)
    t2 = np.zeros(n, np.uint64)
    ms = ctypes.c_float(0)
    print(f"case {name} share={share} batch={batch}", flush=True)
    assert HAST_LIB.ref_hast_run(n, len(rec), rec, share, batch, labels, t1, t2,
                                 ctypes.byref(ms)) == 0
    want = O.bfs_labels(n, u, v)
    assert np.array_equal(labels, want)
    hooked = t2 != np.uint64(0xFFFFFFFFFFFFFFFF)
    assert int(hooked.sum()) == n - len(np.unique(want))
    # the merged forest's records are input edges that span the components
    f = t2[hooked]
    fu = (f >> np.uint64(32)).astype(np.uint32)
    fv = (f & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(O.bfs_labels(n, fu, fv), want)
    assert np.isin(f, rec).all()


def _new_boruvka_lib(H):
    H.harness_new_boruvka.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    H.harness_graph_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32),
                                     ctypes.POINTER(ctypes.c_uint64)]
    H.harness_mmap_populate.argtypes = [ctypes.c_char_p, _u32p, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_int]
    return H


def _mtx(tmp_path, n, a, b, w=None):
    """A MatrixMarket file with 1-based pairs (a, b) and optional weights."""
    p = tmp_path / "g.mtx"
    lines = [f"{x} {y}" + ("" if w is None else f" {z}")
             for x, y, z in zip(a, b, w if w is not None else a)]
    p.write_text("%%MatrixMarket matrix coordinate integer general\n% test\n"
                 f"{n} {n} {len(a)}\n" + "\n".join(lines) + "\n")
    return str(p)


@pytest.mark.parametrize("mode,chunk,nchunks,debug", [
    (0, 0, 0, 0), (1, 7000, 0, 0), (1, 0, 5, 0), (2, 0, 3, 0), (2, 0, 1, 0),
    (2, 9000, 0, 0), (3, 0, 0, 0), (4, 0, 0, 0), (0, 0, 0, 1), (1, 0, 4, 1),
    (2, 0, 3, 1), (3, 0, 0, 1)])
@pytest.mark.parametrize("gen", [0, 1])
def test_new_boruvka_matrix_market(H, tmp_path, mode, chunk, nchunks, debug, gen):
    """The reference's MST driver on a MatrixMarket file: the forest written
    to the result file (u+1 v+1 w, ascending input order) equals the oracle's
    Kruskal forest of the graph the reader loaded, with integer weights (many
    ties) or the generated-weight rule, whole or in chunks, on the GPU, from
    host memory (mode 3, managed), on the CPU (mode 4) and in debug mode (the
    streamline paths repeated 10 times)."""
    import phem_native as P
    L = _new_boruvka_lib(H)
    rng = np.random.default_rng(3 + gen)
    n, m = 3000, 40000
    a = rng.integers(1, n + 1, m)
    b = rng.integers(1, n + 1, m)
    w = rng.integers(0, 50, m)
    path = _mtx(tmp_path, n, a, b, None if gen else w)
    g = P.read_mtx(path, generate_weights=bool(gen))
    ids, lab = O.kruskal(g["n"], g["u"], g["v"], g["w"], 0)
    res = str(tmp_path / "forest.txt")
    print(f"new_boruvka mode={mode} chunk={chunk} nchunks={nchunks} gen={gen} "
          f"debug={debug}", flush=True)
    assert L.harness_new_boruvka(path.encode(), res.encode(), chunk, nchunks, gen, mode,
                                 debug) == 0
    got = np.loadtxt(res, dtype=np.uint64, ndmin=2)
    want = np.stack([g["u"][ids] + 1, g["v"][ids] + 1, g["w"][ids]], axis=1).astype(np.uint64)
    assert got.shape == want.shape and np.array_equal(got, want)
    # the header reader and the array reader agree with the library's reader
    nn, mm = ctypes.c_uint32(0), ctypes.c_uint64(0)
    L.harness_graph_info(path.encode(), ctypes.byref(nn), ctypes.byref(mm))
    assert (nn.value, mm.value) == (n, m)
    aos = np.zeros(3 * m, np.uint32)
    L.harness_mmap_populate(path.encode(), aos, m, n, gen)
    assert np.array_equal(aos.reshape(-1, 3), np.stack([g["u"], g["v"], g["w"]], axis=1))


@pytest.mark.parametrize("share", [0.05, 2.0])
def test_new_boruvka_heterogeneous_host_share(H, tmp_path, share):
    """The heterogeneous mode with a host share set (the default gives the
    host none): the host reduces part (0.05) or all (2.0) of the last chunk,
    and the result file still equals the oracle's forest."""
    import phem_native as P
    L = _new_boruvka_lib(H)
    rng = np.random.default_rng(8)
    n, m = 2000, 60000
    a = rng.integers(1, n + 1, m)
    b = rng.integers(1, n + 1, m)
    w = rng.integers(0, 20, m)
    path = _mtx(tmp_path, n, a, b, w)
    g = P.read_mtx(path)
    ids, _ = O.kruskal(g["n"], g["u"], g["v"], g["w"], 0)
    res = str(tmp_path / "forest.txt")
    old = P.get_tuning()["cpu_share"]
    try:
        P.set_tuning(cpu_share=share)
        print(f"new_boruvka heterogeneous cpu_share={share}", flush=True)
        assert L.harness_new_boruvka(path.encode(), res.encode(), 0, 3, 0, 2, 0) == 0
    finally:
        P.set_tuning(cpu_share=old)
    got = np.loadtxt(res, dtype=np.uint64, ndmin=2)
    want = np.stack([g["u"][ids] + 1, g["v"][ids] + 1, g["w"][ids]], axis=1).astype(np.uint64)
    assert np.array_equal(got, want)


def test_device_boruvka_after_workspace_release(H):
    """device_boruvka still gives the oracle's forest after its cached
    workspace is released, and the release is repeatable."""
    n, u, v, w, wmode = cases.make("rmat_small")
    aos = np.ascontiguousarray(np.stack([u, v, w], axis=1).reshape(-1), np.uint32)
    ids, rec, lab = expected_records(n, u, v, w, 0)
    for _ in range(2):
        H.harness_release_workspace()
        out = np.zeros(3 * n, np.uint32)
        labels = np.zeros(n, np.uint32)
        filled = np.array([0], np.uint32)
        m_after = np.array([7], np.uint32)
        assert H.harness_device_boruvka(aos, n, len(u), out, labels, filled, m_after) == 0
        assert filled[0] == len(ids)
        assert np.array_equal(out[: 3 * len(ids)].reshape(-1, 3), rec)
        assert np.array_equal(labels, lab)
