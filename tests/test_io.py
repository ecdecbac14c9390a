"""Graph file readers (host code of libphem.so, no GPU): hand-computed
fixtures for each format, the reference's generated-weight rule as a closed
form, a large file parsed by several threads against numpy, error paths."""
import struct

import numpy as np
import pytest

import phem_native as P


def write(tmp_path, name, text, mode="w"):
    p = tmp_path / name
    with open(p, mode) as f:
        f.write(text)
    return str(p)


def test_mtx_fixture(tmp_path):
    p = write(tmp_path, "g.mtx",
              "%%MatrixMarket matrix coordinate integer general\n"
              "% a comment\n"
              "5 5 6\n"
              "1 2 7\n"
              "3 1 4\r\n"
              " 2   5 9\n"
              "4 4 1\n"
              "\n"
              "5\t3 2\n"
              "2 3\n"
              "9 9 9\n")
    g = P.read_mtx(p)
    assert (g["n"], g["m"], g["wmode"]) == (5, 6, P.W_U32)
    assert g["u"].tolist() == [0, 0, 1, 3, 2, 1]
    assert g["v"].tolist() == [1, 2, 4, 3, 4, 2]
    assert g["w"].tolist() == [7, 4, 9, 1, 2, 0]  # missing weight reads as 0


def test_mtx_generated_weight_rule(tmp_path):
    n = 100000
    pairs = [(70000, 80000), (1, 1), (3, 2), (100000, 99999), (65536, 65537)]
    p = write(tmp_path, "g.mtx", f"% x\n{n} {n} {len(pairs)}\n" +
              "".join(f"{a} {b} 5\n" for a, b in pairs))
    g = P.read_mtx(p, generate_weights=True)
    # w = ((u - 1)(v - 1) mod n) + 1 with 32-bit unsigned wraparound of the product
    want = [((((a - 1) * (b - 1)) & 0xFFFFFFFF) % n) + 1 for a, b in pairs]
    assert g["w"].tolist() == want
    assert want[0] != (((70000 - 1) * (80000 - 1)) % n) + 1  # the wraparound matters


def test_mtx_large_parallel(tmp_path):
    rng = np.random.default_rng(4)
    m, n = 300000, 50000
    a = rng.integers(1, n + 1, m)
    b = rng.integers(1, n + 1, m)
    w = rng.integers(0, 2**32, m, dtype=np.uint64)
    body = "\n".join(f"{x} {y} {z}" for x, y, z in zip(a, b, w))
    p = write(tmp_path, "big.mtx", f"%%MatrixMarket\n{n} {n} {m}\n{body}\n")
    g = P.read_mtx(p)
    assert g["m"] == m
    assert np.array_equal(g["u"], (np.minimum(a, b) - 1).astype(np.uint32))
    assert np.array_equal(g["v"], (np.maximum(a, b) - 1).astype(np.uint32))
    assert np.array_equal(g["w"], w.astype(np.uint32))


def test_edge_list_fixture(tmp_path):
    p = write(tmp_path, "g.edges", "4 6\n0 1\n1 0\n1 2\n2 1\n3 2\n2 3\n")
    g = P.read_edge_list(p)
    assert (g["n"], g["m"]) == (4, 3)
    assert g["edges"].tolist() == [(0 << 32) | 1, (1 << 32) | 2, (2 << 32) | 3]


def test_ecl_fixture(tmp_path):
    off = [0, 2, 4, 5, 6]
    nb = [1, 2, 0, 3, 0, 1]
    data = struct.pack("<qq", 4, 6) + struct.pack("<5q", *off) + struct.pack("<6i", *nb)
    p = write(tmp_path, "g.egr", data, "wb")
    g = P.read_ecl(p)
    assert (g["n"], g["m"]) == (4, 3)
    assert g["edges"].tolist() == [(0 << 32) | 1, (0 << 32) | 2, (1 << 32) | 3]
    bad = write(tmp_path, "bad.egr", data[:-4], "wb")
    with pytest.raises(P.PhemError) as e:
        P.read_ecl(bad)
    assert e.value.code == P.EINVAL
    # E = 2^62 makes 4 E wrap to 0 in 64 bits: a file of header and offsets
    # only must still be refused (its offsets up to E would read past it)
    huge = 1 << 62
    wrap = struct.pack("<qq", 4, huge) + struct.pack("<5q", 0, 1, 2, 3, huge)
    for name, d in (("wrap.egr", wrap), ("wrap1.egr", wrap + b"\0\0")):
        with pytest.raises(P.PhemError) as e:
            P.read_ecl(write(tmp_path, name, d, "wb"))
        assert e.value.code == P.EINVAL


def test_pbbs_weighted_fixture(tmp_path):
    p = write(tmp_path, "g.pbbs", "WeightedEdgeArray\n0 1 0.5\n2 1 1e-3\n3 0 7\n")
    g = P.read_pbbs_weighted(p)
    assert (g["n"], g["m"], g["wmode"]) == (4, 3, P.W_F32)
    assert g["u"].tolist() == [0, 2, 3] and g["v"].tolist() == [1, 1, 0]
    want = np.array([0.5, 1e-3, 7.0], dtype=np.float64).astype(np.float32).view(np.uint32)
    assert np.array_equal(g["w"], want)
    with pytest.raises(P.PhemError):
        P.read_pbbs_weighted(write(tmp_path, "x.pbbs", "EdgeArray\n0 1\n"))


def test_weighted_dump_round_trip(tmp_path):
    u = np.array([0, 3, 1], np.uint32)
    v = np.array([2, 1, 1], np.uint32)
    w = np.array([5, 4000000000, 0], np.uint32)
    p = str(tmp_path / "weighted.mtx")
    P.write_mtx_weighted(p, 4, u, v, w)
    text = open(p).read().splitlines()
    assert text[0] == "%%MatrixMarket matrix coordinate integer general"
    assert text[3] == "4 4 6"
    assert text[4:6] == ["1 3 5", "3 1 5"]
    g = P.read_mtx(p)
    assert g["m"] == 6
    assert g["u"].tolist() == [0, 0, 1, 1, 1, 1]
    assert g["v"].tolist() == [2, 2, 3, 3, 1, 1]
    assert g["w"].tolist() == [5, 5, 4000000000, 4000000000, 0, 0]


def test_missing_file():
    with pytest.raises(P.PhemError) as e:
        P.read_mtx("/nonexistent/graph.mtx")
    assert e.value.code == P.ENOTFOUND
