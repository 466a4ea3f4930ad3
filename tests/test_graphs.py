"""Host graph generators and CSR builder (CPU only)."""
import numpy as np
import pytest

from jaccard_amd import graphs


def check_csr(g, symmetric=True):
    g.validate()
    off, ind = g.offsets, g.indices
    rows = np.repeat(np.arange(g.n), np.diff(off))
    assert not np.any(rows == ind), "self loop"
    for u in range(min(g.n, 2000)):
        r = ind[off[u]:off[u + 1]]
        assert np.all(np.diff(r) > 0), "row not strictly increasing"
    if symmetric:
        a = set(zip(rows.tolist(), ind.tolist()))
        assert all((v, u) in a for (u, v) in list(a)[:20000])


def test_rmat16_regression_values():
    g = graphs.rmat(16, 16, seed=1)
    assert g.n == 65536 and g.m == 1819866
    st = g.stats()
    assert st["sum_min_deg"] == 325572946
    assert st["max_deg"] == 9769 and st["isolated"] == 18844
    check_csr(g)


def test_rmat_deterministic_and_seeded():
    a = graphs.rmat(12, 8, seed=5)
    b = graphs.rmat(12, 8, seed=5)
    c = graphs.rmat(12, 8, seed=6)
    assert np.array_equal(a.indices, b.indices) and np.array_equal(a.offsets, b.offsets)
    assert not (a.m == c.m and np.array_equal(a.indices, c.indices))


def test_rmat_non_power_of_two_vertices():
    g = graphs.rmat_n(3000, 10, seed=2)
    assert g.n == 3000
    check_csr(g)


def test_weights_symmetric_and_exact():
    g = graphs.rmat(10, 8, seed=3, weighted=True)
    w = g.weights
    assert w.dtype == np.float32 and np.all(w > 0) and np.all(w <= 1)
    k = w.astype(np.float64) * (1 << 24)
    assert np.array_equal(k, np.round(k))  # multiples of 2^-24
    rows = np.repeat(np.arange(g.n), np.diff(g.offsets))
    d = {(u, v): x for u, v, x in zip(rows.tolist(), g.indices.tolist(), w.tolist())}
    assert all(d[(v, u)] == x for (u, v), x in d.items())


def test_chung_lu_calibrated_mean_degree():
    g = graphs.chung_lu(n=1 << 16, mean_deg=32, max_deg=1 << 13, seed=1)
    assert abs(g.m / g.n - 32) / 32 < 0.01
    assert g.degrees.max() <= 1 << 13
    check_csr(g)


def test_from_edges_dedupe_and_selfloops():
    g = graphs.from_edges(5, [0, 0, 1, 2, 3, 3], [1, 1, 1, 4, 3, 2])
    assert g.offsets.tolist() == [0, 1, 2, 4, 5, 6]
    assert g.indices.tolist() == [1, 0, 3, 4, 2, 2]


def test_from_edges_rejects_out_of_range():
    with pytest.raises(ValueError):
        graphs.from_edges(3, [0, 5], [1, 2])


def test_directed_builder_keeps_direction():
    g = graphs.from_edges(3, [0, 1], [1, 2], symmetrize=False)
    assert g.offsets.tolist() == [0, 1, 2, 2] and g.indices.tolist() == [1, 2]


def test_validate_catches_bad_csr():
    g = graphs.rmat(8, 4, seed=1)
    bad = graphs.CSRGraph(n=g.n, offsets=g.offsets.copy(), indices=g.indices.copy())
    bad.indices[0] = g.n
    with pytest.raises(ValueError):
        bad.validate()


def test_rmat_raw_edges_are_the_generator_draws():
    import ctypes
    from jaccard_amd._native import load_graph_lib
    lib = load_graph_lib()
    scale, ef = 10, 8
    e = ef << scale
    s = np.empty(e, np.int32)
    d = np.empty(e, np.int32)
    assert lib.jcg_rmat_edges(scale, 0, float(ef), graphs.RMAT_A, graphs.RMAT_B, graphs.RMAT_C, 5,
                              s.ctypes.data, d.ctypes.data, e) == e
    assert s.min() >= 0 and s.max() < (1 << scale) and d.max() < (1 << scale)
    want = graphs.rmat(scale, ef, seed=5)
    got = graphs.from_edges(1 << scale, s, d)
    assert np.array_equal(got.offsets, want.offsets) and np.array_equal(got.indices, want.indices)
    assert lib.jcg_rmat_edges(scale, 0, float(ef), graphs.RMAT_A, graphs.RMAT_B, graphs.RMAT_C, 5,
                              s.ctypes.data, d.ctypes.data, e - 1) == -1
