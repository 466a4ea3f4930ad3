"""Graph file formats (.mtx, binary .csr v2) and the command-line tools'
host-side behaviour: round trips, header layout, reference error codes,
reverse-edge handling, file_convert and compare_coords. CPU only."""
import os
import struct
import subprocess

import numpy as np
import pytest

from jaccard_amd import fileio, graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sycl-jaccard_amd", "bin")


def tool(name, *args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([os.path.join(BIN, name), *map(str, args)], capture_output=True, text=True, env=e)


def write_mtx(path, header, rows, n, entries):
    with open(path, "w") as f:
        f.write(header + "\n% a comment\n")
        f.write(f"{n} {n} {len(entries)}\n")
        for e in entries:
            f.write(" ".join(str(x) for x in e) + "\n")


def test_csr_header_layout(tmp_path):
    g = graphs.rmat(7, 8, seed=1, weighted=True)
    p = tmp_path / "g.csr"
    fileio.save(str(p), g)
    raw = p.read_bytes()
    version, n, m, flags = struct.unpack("<qqqQ", raw[:32])
    assert (version, n, m) == (2, g.n, g.m)
    assert flags == fileio.WEIGHTED | fileio.ZERO_INDEXED | fileio.HAS_REVERSE_EDGES
    assert len(raw) == 32 + 4 * (g.n + 1) + 4 * g.m + 4 * g.m
    off = np.frombuffer(raw[32:32 + 4 * (g.n + 1)], np.int32)
    assert np.array_equal(off, g.offsets)


@pytest.mark.parametrize("weighted", [False, True])
def test_csr_round_trip(tmp_path, weighted):
    g = graphs.rmat(9, 8, seed=3, weighted=weighted)
    p = str(tmp_path / "g.csr")
    fileio.save(p, g)
    L = fileio.load(p)
    assert L.weighted == weighted
    assert np.array_equal(L.graph.offsets, g.offsets) and np.array_equal(L.graph.indices, g.indices)
    if weighted:
        assert np.array_equal(L.graph.weights, g.weights)


def test_symmetric_mtx_gets_reverse_edges(tmp_path):
    p = str(tmp_path / "t.mtx")
    # triangle 1-2-3 plus pendant 3-4, one direction each, a self loop and a repeat
    write_mtx(p, "%%MatrixMarket matrix coordinate pattern symmetric", None, 5,
              [(2, 1), (3, 1), (3, 2), (4, 3), (4, 4), (3, 1)])
    L = fileio.load(p)
    g = L.graph
    assert g.n == 5 and not L.directed
    want = graphs.from_edges(5, np.array([1, 2, 2, 3]), np.array([0, 0, 1, 2]))
    assert np.array_equal(g.offsets, want.offsets) and np.array_equal(g.indices, want.indices)


def test_general_mtx_is_directed(tmp_path):
    p = str(tmp_path / "d.mtx")
    write_mtx(p, "%%MatrixMarket matrix coordinate integer general", None, 3, [(1, 2, 5), (2, 3, 7)])
    L = fileio.load(p)
    assert L.directed and L.graph.m == 2
    assert list(L.graph.indices) == [1, 2]
    assert L.weighted and list(L.graph.weights) == [1.0, 1.0]  # the driver reads MTX weights as 1.0
    L2 = fileio.load(p, keep_mtx_weights=True)
    assert list(L2.graph.weights) == [5.0, 7.0]


@pytest.mark.parametrize("header,status", [
    ("%%MatrixMarket vector coordinate real general", 1),
    ("%%MatrixMarket matrix array real general", 2),
    ("%%MatrixMarket matrix coordinate complex general", 3),
])
def test_mtx_errors(tmp_path, header, status):
    p = str(tmp_path / "bad.mtx")
    write_mtx(p, header, None, 2, [(1, 2, 1.0)])
    with pytest.raises(fileio.FileFormatError) as e:
        fileio.load(p)
    assert e.value.status == status


def test_csr_64bit_types_rejected(tmp_path):
    p = tmp_path / "w64.csr"
    p.write_bytes(struct.pack("<qqqQ", 2, 1, 0, fileio.WEIGHTED | fileio.WEIGHT_T64) + struct.pack("<ii", 0, 0))
    with pytest.raises(fileio.FileFormatError) as e:
        fileio.load(str(p))
    assert e.value.status == 3


def test_mtx_output_lower_triangle(tmp_path):
    g = graphs.rmat(8, 8, seed=5)
    vals = np.arange(g.m, dtype=np.float32)
    rows = np.repeat(np.arange(g.n), np.diff(g.offsets))
    # symmetric values: J(u,v) = J(v,u)
    sym = (np.minimum(rows, g.indices) * 1000 + np.maximum(rows, g.indices)).astype(np.float32)
    p = str(tmp_path / "o.mtx")
    fileio.save(p, g, values=sym, remove_reverse=True)
    lines = open(p).read().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate real symmetric"
    n1, n2, cnt = map(int, lines[1].split())
    assert (n1, n2, cnt) == (g.n, g.n, g.m // 2)
    ent = np.array([list(map(float, l.split())) for l in lines[2:]])
    assert (ent[:, 0] > ent[:, 1]).all()  # src > dst
    # column-major order
    order = np.lexsort((ent[:, 0], ent[:, 1]))
    assert np.array_equal(order, np.arange(len(ent)))
    del vals


def test_file_convert_and_compare(tmp_path):
    g = graphs.rmat(8, 8, seed=6, weighted=True)
    src = str(tmp_path / "g.csr")
    fileio.save(src, g)
    # csr (both directions) -> mtx (one direction) -> csr (both directions again)
    r = tool("file_convert", src, tmp_path / "g.mtx", 0)
    assert r.returncode == 0, r.stderr
    r = tool("file_convert", tmp_path / "g.mtx", tmp_path / "h.csr", 1)
    assert r.returncode == 0, r.stderr
    back = fileio.load(str(tmp_path / "h.csr")).graph
    assert np.array_equal(back.offsets, g.offsets) and np.array_equal(back.indices, g.indices)
    np.testing.assert_allclose(back.weights, g.weights, rtol=1e-5)
    # identical files compare clean, in either format pairing
    r = tool("compare_coords", src, tmp_path / "h.csr", 1e-5)
    assert r.returncode == 0, r.stderr
    assert "Inputs successfully read" in r.stdout
    r = tool("compare_coords", tmp_path / "g.mtx", tmp_path / "g.mtx", 0)
    assert r.returncode == 0
    # one perturbed value and one missing entry
    w = g.weights.copy()
    w[3] += 0.5
    fileio.save(str(tmp_path / "p.csr"), g, values=w)
    r = tool("compare_coords", src, tmp_path / "p.csr", 1e-3)
    assert r.returncode == 1 and "differ by more than" in r.stderr
    h = graphs.CSRGraph(n=g.n, offsets=np.r_[g.offsets[:-1], g.offsets[-1] - 1].astype(np.int32) if g.degrees[-1] else g.offsets,
                        indices=g.indices[:-1] if g.degrees[-1] else g.indices, weights=None)
    if g.degrees[-1]:
        fileio.save(str(tmp_path / "q.csr"), h, values=g.weights[:-1])
        r = tool("compare_coords", src, tmp_path / "q.csr", 1e-3)
        assert r.returncode >= 1 and "missing from test" in r.stderr


def test_convert_flags_and_errors(tmp_path):
    g = graphs.rmat(6, 8, seed=7)
    src = str(tmp_path / "g.csr")
    fileio.save(src, g)
    r = tool("file_convert", src, tmp_path / "x.txt", 0)
    assert r.returncode == 2
    r = tool("file_convert", tmp_path / "x.txt", src, 0)
    assert r.returncode == 1
    r = tool("file_convert", src, tmp_path / "g.mtx", 1)  # keep reverse with MTX output: refused with a warning
    assert r.returncode == 0 and "Cannot retain reverseEdges with MTX output" in r.stderr
    r = tool("file_convert", src, tmp_path / "d.csr", 0, env={"CONVERT_FORCE_DROP_WEIGHTS": "1"})
    assert r.returncode == 0


def test_driver_argument_errors(tmp_path):
    g = graphs.rmat(6, 8, seed=8)
    src = str(tmp_path / "g.csr")
    fileio.save(src, g)
    assert tool("jaccard_amd", tmp_path / "g.txt", tmp_path / "o.mtx").returncode == 1
    assert tool("jaccard_amd", src, tmp_path / "o.txt").returncode == 2
    p = tmp_path / "w64.csr"
    p.write_bytes(struct.pack("<qqqQ", 2, 1, 0, fileio.WEIGHTED | fileio.WEIGHT_T64) + struct.pack("<ii", 0, 0))
    r = tool("jaccard_amd", p, tmp_path / "o.mtx")
    assert r.returncode == 3 and "does not match required data types" in r.stderr
