"""GPU parity: the HIP library through its C ABI against the CPU oracle.

Bar: unweighted outputs bit-exact (EC and VC, reference arithmetic
I / ((float)(d_u+d_v) - I)); weighted outputs within 1e-6 relative error of
the FP64 oracle, EC and VC bit-identical to each other.
"""
import numpy as np
import pytest

import oracle
from jaccard_amd import api, graphs

import cases

pytestmark = pytest.mark.gpu

WEIGHTED_RTOL = 1e-6


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bits_equal(got, want, what):
    got, want = np.asarray(got, np.float32), np.asarray(want, np.float32)
    assert got.shape == want.shape, what
    bad = np.nonzero(bits(got) != bits(want))[0]
    assert bad.size == 0, f"{what}: {bad.size} mismatches, first at {bad[:5]}: got {got[bad[:5]]} want {want[bad[:5]]}"


def rel_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    both_nan = np.isnan(got) & np.isnan(want)
    d = np.abs(got - want) / np.maximum(np.abs(want), np.finfo(np.float64).tiny)
    d[both_nan] = 0.0
    d[np.isnan(d)] = np.inf
    return d


def run(g, cuda, weights=None, ec=True, **kw):
    dg = api.DeviceCSR.from_host(g, cuda)
    w = None if weights is None else __import__("torch").from_numpy(np.asarray(weights, np.float32)).to(cuda)
    return api.jaccard(dg, weights=w, edge_centric=ec, **kw).cpu().numpy()


SUITE = cases.small_suite()


@pytest.mark.parametrize("name", list(SUITE))
def test_unweighted_bit_exact(name, cuda):
    g = SUITE[name]
    print(f"[gpu] unweighted {name} n={g.n} m={g.m}", flush=True)
    ref = oracle.ec_unweighted(g)
    assert_bits_equal(ref, oracle.vc_unweighted(g), f"{name}: oracle EC vs VC")
    assert_bits_equal(run(g, cuda, ec=True), ref, f"{name} EC")
    assert_bits_equal(run(g, cuda, ec=False), ref, f"{name} VC")


@pytest.mark.parametrize("name", ["rmat10", "rmat12_ef32", "chunglu_4k", "K130", "two_triangles"])
def test_weighted_vs_fp64(name, cuda):
    g = SUITE[name]
    print(f"[gpu] weighted {name} n={g.n} m={g.m}", flush=True)
    rng = np.random.default_rng(21)
    v = (rng.integers(1, 1 << 24, g.n) / float(1 << 24)).astype(np.float32)
    ref = oracle.weighted_f64(g, v.astype(np.float64))
    ec = run(g, cuda, v, ec=True)
    vc = run(g, cuda, v, ec=False)
    assert rel_err(ec, ref).max() <= WEIGHTED_RTOL
    assert_bits_equal(ec, vc, f"{name} weighted EC vs VC")


@pytest.mark.parametrize("kind", ["dense_span", "wide_span"])
def test_hub_multichunk(kind, cuda):
    """A hub row of degree ~30000: with ids 0..29999 it fits one LDS bitmap;
    spread over 2^21 ids it spans several hash-table chunks in both modes."""
    g = cases.hub_graph() if kind == "dense_span" else cases.sparse_hub_graph()
    print(f"[gpu] hub {kind} n={g.n} m={g.m}", flush=True)
    ref = oracle.ec_unweighted(g)
    assert_bits_equal(run(g, cuda, ec=True), ref, "hub EC")
    assert_bits_equal(run(g, cuda, ec=False), ref, "hub VC")
    dg = api.DeviceCSR.from_host(g, cuda)
    W = api.vertex_weights(dg)
    Wh = W.cpu().numpy()
    np.testing.assert_allclose(Wh, oracle.vertex_weights(g), rtol=1e-7)
    refw = oracle.weighted_f64(g, Wh.astype(np.float64))
    ecw = api.jaccard(dg, weights=W, edge_centric=True).cpu().numpy()
    vcw = api.jaccard(dg, weights=W, edge_centric=False).cpu().numpy()
    assert rel_err(ecw, refw).max() <= WEIGHTED_RTOL
    assert_bits_equal(ecw, vcw, "hub weighted EC vs VC")


def test_bitmap_and_hash_task_lists(cuda):
    """Hub rows on both sides of the bitmap limits: the scheduler puts them on
    different task lists (separate kernel instantiations); every result must
    match the oracle either way."""
    g = cases.span_hubs_graph()
    print(f"[gpu] span hubs n={g.n} m={g.m}", flush=True)
    ref = oracle.ec_unweighted(g)
    assert_bits_equal(run(g, cuda, ec=True), ref, "span hubs EC")
    assert_bits_equal(run(g, cuda, ec=False), ref, "span hubs VC")
    dg = api.DeviceCSR.from_host(g, cuda)
    W = api.vertex_weights(dg)
    refw = oracle.weighted_f64(g, W.cpu().numpy().astype(np.float64))
    ecw = api.jaccard(dg, weights=W, edge_centric=True).cpu().numpy()
    vcw = api.jaccard(dg, weights=W, edge_centric=False).cpu().numpy()
    assert rel_err(ecw, refw).max() <= WEIGHTED_RTOL
    assert_bits_equal(ecw, vcw, "span hubs weighted EC vs VC")


@pytest.mark.parametrize("name", ["rmat12_ef32", "K130", "hub_wide"])
def test_row_load_paths_agree(name, cuda):
    """Rows are streamed with 32-bit buffer offsets (index arrays below 2^30
    entries) or with 64-bit addresses; forcing the latter (wide_index) must
    give the same bits, both equal to the oracle."""
    g = cases.sparse_hub_graph() if name == "hub_wide" else SUITE[name]
    print(f"[gpu] load paths {name} n={g.n} m={g.m}", flush=True)
    dg = api.DeviceCSR.from_host(g, cuda)
    ref = oracle.ec_unweighted(g)
    rng = np.random.default_rng(5)
    v = (rng.integers(1, 1 << 24, g.n) / float(1 << 24)).astype(np.float32)
    W = __import__("torch").from_numpy(v).to(cuda)
    refw = oracle.weighted_f64(g, v.astype(np.float64))
    for ec in (True, False):
        narrow = api.jaccard(dg, edge_centric=ec).cpu().numpy()
        wide = api.jaccard(dg, edge_centric=ec, tuning={"wide_index": 1}).cpu().numpy()
        assert_bits_equal(narrow, ref, f"{name} ec={ec} buffer loads")
        assert_bits_equal(wide, ref, f"{name} ec={ec} 64-bit loads")
        narrow_w = api.jaccard(dg, weights=W, edge_centric=ec).cpu().numpy()
        wide_w = api.jaccard(dg, weights=W, edge_centric=ec, tuning={"wide_index": 1}).cpu().numpy()
        assert_bits_equal(wide_w, narrow_w, f"{name} ec={ec} weighted, load paths")
        assert rel_err(narrow_w, refw).max() <= WEIGHTED_RTOL


@pytest.mark.parametrize("ec", [True, False])
def test_asymmetric_fixup(ec, cuda):
    g = cases.random_directed()
    print(f"[gpu] asymmetric n={g.n} m={g.m} ec={ec}", flush=True)
    dg = api.DeviceCSR.from_host(g, cuda)
    assert not api.is_symmetric(dg)
    assert_bits_equal(run(g, cuda, ec=ec), oracle.ec_unweighted(g), "asymmetric")
    rng = np.random.default_rng(2)
    v = rng.random(g.n).astype(np.float32) + 0.5
    ref = oracle.weighted_f64(g, v.astype(np.float64))
    assert rel_err(run(g, cuda, v, ec=ec), ref).max() <= WEIGHTED_RTOL
    # a result view that is not 16-byte aligned (the fix-up pass scans it)
    import torch
    buf = torch.empty(g.m + 1, dtype=torch.float32, device=cuda)
    out = buf[1:]
    assert out.data_ptr() % 16 != 0
    api.jaccard(dg, out, edge_centric=ec)
    assert_bits_equal(out.cpu().numpy(), oracle.ec_unweighted(g), "asymmetric, unaligned result")


def test_self_loops(cuda):
    g = cases.with_self_loops(graphs.rmat(9, 8, seed=4))
    print(f"[gpu] self loops n={g.n} m={g.m}", flush=True)
    ref = oracle.ec_unweighted(g)
    assert_bits_equal(run(g, cuda, ec=True), ref, "self loops EC")
    assert_bits_equal(run(g, cuda, ec=False), ref, "self loops VC")


def test_edge_cases(cuda):
    import torch
    # no vertices
    e = graphs.CSRGraph(n=0, offsets=np.zeros(1, np.int32), indices=np.zeros(0, np.int32))
    assert run(e, cuda).size == 0
    # isolated vertices only
    iso = graphs.CSRGraph(n=10, offsets=np.zeros(11, np.int32), indices=np.zeros(0, np.int32))
    assert run(iso, cuda).size == 0
    # a single undirected edge: J = 0
    one = graphs.from_edges(3, [0], [2])
    assert_bits_equal(run(one, cuda), np.zeros(2, np.float32), "single edge")
    # workspace variant and symmetric flag give the same bits
    g = SUITE["rmat10"]
    dg = api.DeviceCSR.from_host(g, cuda)
    nbytes = api.workspace_size(g.n, g.m, False, True)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=cuda)
    a = api.jaccard(dg, workspace=ws).cpu().numpy()
    b = api.jaccard(dg, flags=api.FLAG_ASSUME_SYMMETRIC).cpu().numpy()
    assert_bits_equal(a, oracle.ec_unweighted(g), "workspace")
    assert_bits_equal(b, oracle.ec_unweighted(g), "assume symmetric")
    small = torch.empty(8, dtype=torch.uint8, device=cuda)
    with pytest.raises(api.JaccardError) as ei:
        api.jaccard(dg, workspace=small)
    assert ei.value.code == api.ERR_WORKSPACE


def test_pairs(cuda):
    import torch
    g = SUITE["rmat12_ef32"]
    rng = np.random.default_rng(8)
    k = rng.integers(0, g.m, 3000)
    rows = np.searchsorted(g.offsets, k, side="right") - 1
    first = np.concatenate([rows, rng.integers(0, g.n, 3000), [5, 0]]).astype(np.int32)
    second = np.concatenate([g.indices[k], rng.integers(0, g.n, 3000), [5, 0]]).astype(np.int32)
    dg = api.DeviceCSR.from_host(g, cuda)
    f, s = torch.from_numpy(first).to(cuda), torch.from_numpy(second).to(cuda)
    out = api.jaccard_list(dg, f, s).cpu().numpy()
    ref = oracle.pairs_unweighted(g, first, second)
    assert_bits_equal(out, ref, "pairs unweighted")  # NaN payloads of 0/0 compare bitwise too
    full = run(g, cuda)
    assert_bits_equal(out[:3000], full[k], "pair value == full-graph value")
    v = (rng.random(g.n) + 0.1).astype(np.float32)
    outw = api.jaccard_list(dg, f, s, weights=torch.from_numpy(v).to(cuda)).cpu().numpy()
    refw = oracle.pairs_weighted_f64(g, v.astype(np.float64), first, second)
    assert rel_err(outw, refw).max() <= WEIGHTED_RTOL
    fullw = run(g, cuda, v)
    assert_bits_equal(outw[:3000], fullw[k], "weighted pair value == full-graph value")
    # out-of-range ids give NaN, not a fault
    bad = torch.tensor([-1, g.n, 0], dtype=torch.int32, device=cuda)
    ok = torch.tensor([0, 0, g.n + 5], dtype=torch.int32, device=cuda)
    assert np.isnan(api.jaccard_list(dg, bad, ok).cpu().numpy()).all()


def test_length_class_boundaries(cuda):
    """Owned rows on both sides of every streaming class limit (8, 16, 32, 64
    and 512 entries per chunk) under an owner whose keys fit one table chunk
    and under one that needs several (30,000 entries): EC and VC bit-exact
    unweighted, weighted within tolerance."""
    rng = np.random.default_rng(11)
    n = 60000
    lens = [8, 9, 16, 17, 32, 33, 64, 65, 511, 512, 513, 514, 1025, 1026, 1537, 1540, 2000]
    L = len(lens)
    u1, u2 = 0, L + 1          # owners: 2,500 and 30,000 entries
    src, dst = [], []
    for i, d in enumerate(lens):
        x = 1 + i
        nb = rng.choice(np.arange(L + 2, n), size=d - 2, replace=False)
        src += [x] * d
        dst += [u1, u2] + nb.tolist()
    for u, k in ((u1, 2500), (u2, 30000)):
        nb = rng.choice(np.arange(L + 2, n), size=k, replace=False)
        src += [u] * k
        dst += nb.tolist()
    g = graphs.from_edges(n, np.array(src), np.array(dst))
    deg = np.diff(g.offsets)
    assert all(deg[1 + i] == d for i, d in enumerate(lens))
    print(f"[gpu] length classes n={g.n} m={g.m} max_deg={deg.max()}", flush=True)
    ref = oracle.ec_unweighted(g)
    assert_bits_equal(run(g, cuda, ec=True), ref, "class limits EC")
    assert_bits_equal(run(g, cuda, ec=False), ref, "class limits VC")
    v = (rng.integers(1, 1 << 24, g.n) / float(1 << 24)).astype(np.float32)
    refw = oracle.weighted_f64(g, v.astype(np.float64))
    for ec in (True, False):
        assert rel_err(run(g, cuda, v, ec=ec), refw).max() <= WEIGHTED_RTOL


def test_row_sums_at_length_boundaries(cuda):
    """Row sums route rows by length (4 lanes up to 64 entries, 16 per step;
    a whole wave up to 2048; a workgroup beyond); rows on both sides of each
    limit must give the exact sums: vertex weights (edge weights of the row)
    and the weighted RowSum (vertex weights gathered over the row's
    columns)."""
    import torch
    rng = np.random.default_rng(9)
    degs = [0, 1, 7, 8, 9, 15, 16, 17, 32, 33, 63, 64, 65, 66, 2047, 2048, 2049, 5000, 3, 0, 129, 300, 1000]
    n = 6000
    rows = []
    for r, d in enumerate(degs):
        cols = rng.choice(np.setdiff1d(np.arange(n), [r]), size=d, replace=False)
        rows.append(np.sort(cols).astype(np.int32))
    rows += [np.zeros(0, np.int32)] * (n - len(degs))
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([r.size for r in rows])
    ind = np.concatenate(rows).astype(np.int32)
    w = (rng.integers(1, 1 << 24, ind.size) / float(1 << 24)).astype(np.float32)
    g = graphs.CSRGraph(n=n, offsets=off.astype(np.int32), indices=ind, weights=w, name="row_lengths")
    print(f"[gpu] row sums n={g.n} m={g.m}", flush=True)
    dg = api.DeviceCSR.from_host(g, cuda)
    W = api.vertex_weights(dg).cpu().numpy()
    assert_bits_equal(W, oracle.vertex_weights(g).astype(np.float32), "vertex weights")
    v = (rng.integers(1, 1 << 20, n) / float(1 << 20)).astype(np.float32)
    work = torch.empty(n, dtype=torch.float32, device=cuda)
    api.rowsum(dg, work, v=torch.from_numpy(v).to(cuda))
    r_of = np.repeat(np.arange(n), np.diff(off))
    S = np.bincount(r_of, weights=v[ind].astype(np.float64), minlength=n)
    assert_bits_equal(work.cpu().numpy(), S.astype(np.float32), "weighted rowsum")


def test_kernel_level_entry_points(cuda):
    import torch
    g = SUITE["rmat10"]
    dg = api.DeviceCSR.from_host(g, cuda)
    n, m = g.n, g.m
    deg = np.diff(g.offsets).astype(np.float32)
    # RowSum unweighted: degrees
    work = torch.empty(n, dtype=torch.float32, device=cuda)
    api.rowsum(dg, work)
    assert_bits_equal(work.cpu().numpy(), deg, "rowsum unweighted")
    # RowSum weighted: sum of v over the row's columns
    rng = np.random.default_rng(3)
    v = (rng.integers(1, 1 << 20, n) / float(1 << 20)).astype(np.float32)
    vt = torch.from_numpy(v).to(cuda)
    worw = torch.empty(n, dtype=torch.float32, device=cuda)
    api.rowsum(dg, worw, v=vt)
    rows = np.repeat(np.arange(n), np.diff(g.offsets))
    S = np.bincount(rows, weights=v[g.indices].astype(np.float64), minlength=n)
    assert_bits_equal(worw.cpu().numpy(), S.astype(np.float32), "rowsum weighted (exact sums)")
    # Fill + Is + Jw == reference VC pipeline (unweighted, fp32 buffers)
    wi = torch.empty(m, dtype=torch.float32, device=cuda)
    ws = torch.empty(m, dtype=torch.float32, device=cuda)
    api.fill(wi, 0.0)
    api.intersect(dg, work, wi, ws)
    wj = torch.empty(m, dtype=torch.float32, device=cuda)
    api.jw(wi, ws, wj)
    assert_bits_equal(wj.cpu().numpy(), oracle.vc_unweighted(g), "fill/is/jw pipeline")
    assert_bits_equal(ws.cpu().numpy(), deg[rows] + deg[g.indices], "weight_s")
    # Is accumulates into weight_i, as the reference's atomics do
    wi2 = torch.full((m,), 1.0, dtype=torch.float32, device=cuda)
    api.intersect(dg, work, wi2, ws)
    assert_bits_equal(wi2.cpu().numpy(), wi.cpu().numpy() + np.float32(1.0), "Is accumulates")
    # EC scan + EC kernel
    dest = torch.empty(m, dtype=torch.int32, device=cuda)
    api.ec_scan(dg, dest, wj)
    assert np.array_equal(dest.cpu().numpy(), rows.astype(np.int32))
    assert not wj.cpu().numpy().any()
    api.ec_unweighted(dg, dest, wj)
    assert_bits_equal(wj.cpu().numpy(), oracle.ec_unweighted(g), "ec_unweighted")
    # non zero pairs count (reference threshold 1e-5)
    assert api.nonzero_count(wj) == oracle.nonzero_count(oracle.ec_unweighted(g))
    # IsPairs on the stored edges equals Is
    f = torch.from_numpy(rows.astype(np.int32)).to(cuda)
    s = torch.from_numpy(g.indices).to(cuda)
    pi = torch.zeros(m, dtype=torch.float32, device=cuda)
    ps = torch.empty(m, dtype=torch.float32, device=cuda)
    api.intersect_pairs(dg, f, s, work, pi, ps)
    assert_bits_equal(pi.cpu().numpy(), wi.cpu().numpy(), "IsPairs == Is")
    assert_bits_equal(ps.cpu().numpy(), ws.cpu().numpy(), "IsPairs weight_s")


def test_validate_and_symmetry(cuda):
    import torch
    g = SUITE["rmat10"]
    dg = api.DeviceCSR.from_host(g, cuda)
    assert api.validate_csr(dg) == 0
    assert api.is_symmetric(dg)
    ind = g.indices.copy()
    r = int(np.argmax(np.diff(g.offsets)))
    a, b = g.offsets[r], g.offsets[r + 1]
    ind[a:b] = ind[a:b][::-1]
    bad = api.DeviceCSR(g.n, dg.offsets, torch.from_numpy(ind).to(cuda))
    assert api.validate_csr(bad) == 4
    ind2 = g.indices.copy()
    ind2[3] = g.n + 7
    assert api.validate_csr(api.DeviceCSR(g.n, dg.offsets, torch.from_numpy(ind2).to(cuda))) & 2
    off = g.offsets.copy()
    off[5], off[6] = off[6], off[5] - 1
    assert api.validate_csr(api.DeviceCSR(g.n, torch.from_numpy(off).to(cuda), dg.indices)) & 1


def test_weighted_deterministic(cuda):
    g = SUITE["chunglu_4k"]
    rng = np.random.default_rng(9)
    v = rng.random(g.n).astype(np.float32)
    a = run(g, cuda, v)
    b = run(g, cuda, v)
    assert_bits_equal(a, b, "repeat")


def test_rmat16_c1(cuda):
    """The CPU-runnable workload: RMAT scale 16, edge factor 16, unweighted."""
    g = graphs.rmat(16, 16, seed=1)
    print(f"[gpu] rmat16 n={g.n} m={g.m}", flush=True)
    ref = oracle.ec_unweighted(g)
    assert_bits_equal(run(g, cuda, ec=True), ref, "rmat16 EC")
    assert_bits_equal(run(g, cuda, ec=False), ref, "rmat16 VC")


GOLDEN = sorted(__import__("glob").glob(__import__("os").path.join(__import__("os").path.dirname(__file__),
                                                                     "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=[p.rsplit("/", 1)[-1][:-4] for p in GOLDEN])
def test_golden_fixtures(path, cuda):
    """The committed scipy/networkx fixtures, straight through the C ABI."""
    import torch
    z = np.load(path)
    g = graphs.CSRGraph(n=int(z["n"]), offsets=z["offsets"], indices=z["indices"],
                        weights=z["edge_w"] if "edge_w" in z.files else None)
    print(f"[gpu] golden {path.rsplit('/', 1)[-1]} n={g.n} m={g.m}", flush=True)
    for ec in (True, False):
        assert_bits_equal(run(g, cuda, ec=ec), z["expected_unweighted"], f"golden unweighted ec={ec}")
        got = run(g, cuda, z["vertex_w"], ec=ec)
        assert rel_err(got, z["expected_weighted"]).max() <= WEIGHTED_RTOL
    dg = api.DeviceCSR.from_host(g, cuda)
    f = torch.from_numpy(z["first"]).to(cuda)
    s = torch.from_numpy(z["second"]).to(cuda)
    assert_bits_equal(api.jaccard_list(dg, f, s).cpu().numpy(), z["expected_pairs_unweighted"], "golden pairs")
    pw = api.jaccard_list(dg, f, s, weights=torch.from_numpy(z["vertex_w"]).to(cuda)).cpu().numpy()
    assert rel_err(pw, z["expected_pairs_weighted"]).max() <= WEIGHTED_RTOL
    if g.weights is not None:
        W = api.vertex_weights(dg).cpu().numpy()
        np.testing.assert_allclose(W, z["vertex_w_expected"], rtol=6e-8)
