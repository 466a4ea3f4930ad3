"""Weighted paths on the GPU beyond the (0,1] workloads.

* Double-double sums: vertex weights mixing 2^8 and 2^-40 magnitudes span
  about 87 bits, more than plain fp64 holds exactly, so the library must add
  them in double-double (jc_weighted_sum_mode reports the choice the kernels
  make for the same graph and weights). Edge-centric, vertex-centric and pair
  results are checked against the FP64 oracle (1e-6), EC against VC and pair
  against full-graph values bit for bit, on graphs that reach every kernel
  path: one-wave rows, hash-table tasks with several chunks, bitmap tasks.
* Degenerate weights: all zero, or non-zero only on isolated vertices (every
  sum 0): plain fp64 is exact and every result is 0/0 = NaN.
* The kernel-level weighted entries (Jaccard_IsKernel<true>,
  jaccard.cpp:402-458; Jaccard_IsPairsKernel<true>, jaccard.cpp:542-605):
  weight_i gains the vertex-weight sum over the intersection (scipy
  A diag(v) A^T, fp64), weight_s = work[u] + work[v] in fp32.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from jaccard_amd import api

import cases

pytestmark = pytest.mark.gpu

WEIGHTED_RTOL = 1e-6


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def rel_err(got, want):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    both_nan = np.isnan(got) & np.isnan(want)
    d = np.abs(got - want) / np.maximum(np.abs(want), np.finfo(np.float64).tiny)
    d[both_nan] = 0.0
    d[np.isnan(d)] = np.inf
    return d


def wide_span_weights(n, seed):
    """fp32 weights 2^8 (1 + j 2^-23) and 2^-40 (1 + j 2^-23): finest bit
    2^-63, row sums up to about 2^24, a span of about 87 bits."""
    rng = np.random.default_rng(seed)
    frac = 1.0 + rng.integers(0, 1 << 23, n) * 2.0 ** -23
    big = rng.random(n) < 0.5
    return np.where(big, 256.0 * frac, 2.0 ** -40 * frac).astype(np.float32)


GRAPHS = {
    "rmat12_ef32": lambda: cases.small_suite()["rmat12_ef32"],
    "hub_bitmap": cases.hub_graph,           # hub row on the bitmap task list
    "hub_hash_chunks": cases.sparse_hub_graph,  # hub row over several hash-table chunks
}


@pytest.mark.parametrize("name", list(GRAPHS))
def test_double_double_sums(name, cuda):
    import torch
    g = GRAPHS[name]()
    print(f"[gpu] double-double {name} n={g.n} m={g.m}", flush=True)
    v = wide_span_weights(g.n, seed=31)
    dg = api.DeviceCSR.from_host(g, cuda)
    vt = torch.from_numpy(v).to(cuda)
    assert api.weighted_sum_mode(dg, vt) == api.SUM_DOUBLE_DOUBLE
    grid = (np.random.default_rng(2).integers(1, 1 << 24, g.n) / float(1 << 24)).astype(np.float32)
    assert api.weighted_sum_mode(dg, torch.from_numpy(grid).to(cuda)) == api.SUM_FP64
    ref = oracle.weighted_f64(g, v.astype(np.float64))
    ec = api.jaccard(dg, weights=vt, edge_centric=True).cpu().numpy()
    vc = api.jaccard(dg, weights=vt, edge_centric=False).cpu().numpy()
    err = rel_err(ec, ref)
    print(f"[gpu]   EC max rel err {err.max():.3e}", flush=True)
    assert err.max() <= WEIGHTED_RTOL
    assert np.array_equal(bits(ec), bits(vc)), "weighted EC and VC differ"
    rng = np.random.default_rng(4)
    k = rng.integers(0, g.m, 4000)
    rows = (np.searchsorted(g.offsets, k, side="right") - 1).astype(np.int32)
    cols = g.indices[k].astype(np.int32)
    f, s = torch.from_numpy(rows).to(cuda), torch.from_numpy(cols).to(cuda)
    pr = api.jaccard_list(dg, f, s, weights=vt).cpu().numpy()
    assert np.array_equal(bits(pr), bits(ec[k])), "pair value differs from the full-graph value"
    assert rel_err(pr, oracle.pairs_weighted_f64(g, v.astype(np.float64), rows, cols)).max() <= WEIGHTED_RTOL


@pytest.mark.parametrize("kind", ["all_zero", "zero_row_sums"])
def test_degenerate_weights(kind, cuda):
    import torch
    g = cases.small_suite()["rmat10"]
    deg = np.diff(g.offsets)
    v = np.zeros(g.n, np.float32)
    if kind == "zero_row_sums":
        iso = np.nonzero(deg == 0)[0]
        assert iso.size > 0
        v[iso] = 0.75  # weights only where no row can see them
    print(f"[gpu] degenerate weights {kind} n={g.n} m={g.m}", flush=True)
    dg = api.DeviceCSR.from_host(g, cuda)
    vt = torch.from_numpy(v).to(cuda)
    assert api.weighted_sum_mode(dg, vt) == api.SUM_FP64
    for ec in (True, False):
        out = api.jaccard(dg, weights=vt, edge_centric=ec).cpu().numpy()
        assert np.isnan(out).all(), f"0/0 expected everywhere (ec={ec})"


def _intersection_weights(g, v, first, second):
    A = sp.csr_matrix((np.ones(g.m), g.indices, g.offsets), shape=(g.n, g.n))
    M = (A @ sp.diags(v.astype(np.float64)) @ A.T).tocsr()
    return np.asarray(M[first, second]).ravel()


@pytest.mark.parametrize("weights", ["grid", "wide_span"])
def test_kernel_level_weighted_intersection(weights, cuda):
    import torch
    g = cases.small_suite()["rmat10"]
    print(f"[gpu] weighted Is/IsPairs ({weights}) n={g.n} m={g.m}", flush=True)
    rng = np.random.default_rng(12)
    v = ((rng.integers(1, 1 << 20, g.n) / float(1 << 20)).astype(np.float32) if weights == "grid"
         else wide_span_weights(g.n, seed=13))
    dg = api.DeviceCSR.from_host(g, cuda)
    vt = torch.from_numpy(v).to(cuda)
    work = torch.empty(g.n, dtype=torch.float32, device=cuda)
    api.rowsum(dg, work, v=vt)
    wk = work.cpu().numpy()
    rows = (np.repeat(np.arange(g.n), np.diff(g.offsets))).astype(np.int32)
    cols = g.indices.astype(np.int32)
    Iw = _intersection_weights(g, v, rows, cols)
    # Is<true> adds into weight_i (like the reference's atomics), sets weight_s
    wi = torch.full((g.m,), 0.5, dtype=torch.float32, device=cuda)
    ws = torch.empty(g.m, dtype=torch.float32, device=cuda)
    api.intersect(dg, work, wi, ws, v=vt)
    want_i = np.float32(0.5) + Iw.astype(np.float32)
    assert rel_err(wi.cpu().numpy(), want_i).max() <= WEIGHTED_RTOL
    assert np.array_equal(bits(ws.cpu().numpy()), bits(wk[rows] + wk[cols])), "weight_s"
    # IsPairs<true> on stored edges and random pairs
    pf = np.concatenate([rows[::3], rng.integers(0, g.n, 2000)]).astype(np.int32)
    ps_ = np.concatenate([cols[::3], rng.integers(0, g.n, 2000)]).astype(np.int32)
    f, s = torch.from_numpy(pf).to(cuda), torch.from_numpy(ps_).to(cuda)
    pi = torch.full((pf.size,), 0.25, dtype=torch.float32, device=cuda)
    pws = torch.empty(pf.size, dtype=torch.float32, device=cuda)
    api.intersect_pairs(dg, f, s, work, pi, pws, v=vt)
    want_p = np.float32(0.25) + _intersection_weights(g, v, pf, ps_).astype(np.float32)
    assert rel_err(pi.cpu().numpy(), want_p).max() <= WEIGHTED_RTOL
    assert np.array_equal(bits(pws.cpu().numpy()), bits(wk[pf] + wk[ps_])), "pairs weight_s"
    # stored edges: the pair entry adds the same value as the full-graph entry
    got_full = wi.cpu().numpy()[::3] - np.float32(0.5)
    got_pair = pi.cpu().numpy()[: rows[::3].size] - np.float32(0.25)
    assert rel_err(got_pair, got_full).max() <= WEIGHTED_RTOL
