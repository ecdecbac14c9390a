"""The C ABI library loads and exports every function include/phem.h declares
(CPU: no kernel is launched)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "phem-graph_amd", "libphem.so")


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(phem_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libphem.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libphem.so not built")
def test_workspace_query_without_device():
    lib = ctypes.CDLL(LIB)
    b = ctypes.c_uint64(0)
    assert lib.phem_mst_workspace_bytes(ctypes.c_uint32(1 << 24), ctypes.c_uint64(1 << 28),
                                        ctypes.byref(b)) == 0
    # two 16-byte edge buffers dominate
    assert 2 * 16 * (1 << 28) <= b.value < 2 * 16 * (1 << 28) + (1 << 30)
    assert lib.phem_mst_workspace_bytes(ctypes.c_uint32(10), ctypes.c_uint64(1 << 40),
                                        ctypes.byref(b)) == -1
    lib.phem_last_error.restype = ctypes.c_char_p
    assert b"beyond limits" in lib.phem_last_error()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libphem.so not built")
def test_stream_footprint_query_without_device():
    """phem_stream_bytes (what the server admits a resident stream by) adds up
    the stream's allocations: two [forest | batch] buffers of 20 B per slot,
    the solve's workspace for n + max_batch edges, 24 B per vertex and 8 B
    per batch edge; bad sizes are refused before any device call."""
    lib = ctypes.CDLL(LIB)
    n, mb = 1 << 20, 1 << 16
    b, ws = ctypes.c_uint64(0), ctypes.c_uint64(0)
    assert lib.phem_stream_bytes(ctypes.c_uint32(n), ctypes.c_uint64(mb), ctypes.byref(b)) == 0
    assert lib.phem_mst_workspace_bytes(ctypes.c_uint32(n), ctypes.c_uint64(n + mb),
                                        ctypes.byref(ws)) == 0
    assert b.value == 40 * (n + mb) + ws.value + 20 * n + 8 * mb
    for bad in ((0, 1), (1, 0), (1, 1 << 40)):
        assert lib.phem_stream_bytes(ctypes.c_uint32(bad[0]), ctypes.c_uint64(bad[1]),
                                     ctypes.byref(b)) == -1
    assert lib.phem_server_stream_open(None, ctypes.c_uint32(1), ctypes.c_uint64(1),
                                       ctypes.byref(b)) == -1
    assert lib.phem_server_stream_close(None, ctypes.c_uint64(1)) == -1


@pytest.mark.skipif(not os.path.exists(LIB), reason="libphem.so not built")
def test_noalloc_refuses_graphs_beyond_its_scratch():
    """The size check comes before any device call: a graph whose workspace
    exceeds the 64 MiB static scratch gets PHEM_ENOMEM with a message."""
    lib = ctypes.CDLL(LIB)
    lib.phem_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    lib.phem_mst_solve_noalloc.argtypes = [ctypes.c_uint32, ctypes.c_uint64, vp, vp, vp,
                                           ctypes.c_int, vp, vp, vp, vp, vp]
    fake = vp(0x1000)  # never dereferenced: the call fails before any use
    rc = lib.phem_mst_solve_noalloc(1 << 17, 1 << 21, fake, fake, fake, 1, fake, None,
                                    None, None, None)
    assert rc == -2
    assert b"static scratch" in lib.phem_last_error()
    b = ctypes.c_uint64(0)
    assert lib.phem_mst_workspace_bytes(ctypes.c_uint32(1 << 17), ctypes.c_uint64(1 << 20),
                                        ctypes.byref(b)) == 0
    assert b.value <= 64 << 20
