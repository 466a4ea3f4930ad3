"""GPU tests at the full sizes of the measured workloads.

* RMAT-22 weighted (n = 4,194,304, m = 128,309,750; the headline graph):
  edge-centric and vertex-centric weighted results equal bit for bit over
  every stored edge; J(u,v) == J(v,u) bitwise at every mirror position;
  0 <= J <= 1; a 200,000-edge sample plus every edge of the five largest hubs
  against the FP64 oracle (1e-6, zeros exact); unweighted EC on the same CSR
  bit-exact against the oracle's restatement of the reference loop
  (jaccard.cpp:711-742) on the sample.
* Chung-Lu 2^20 (gamma 2.1, mean degree 32, max 2^17), unweighted: EC and VC
  bit-exact against the oracle over every stored edge.
* Batched: the 4096-graph batch (n log-uniform in [2^8, 2^16], edge factor in
  [4, 32], half weighted) in one union pass equals one plain call per graph
  for every member, bit for bit; 50 members against the oracle.
* Serving: 64 resident RMAT graphs of scales 14-20 (edge factor 16,
  weighted): each request of a mixed stream gives the same bits in the
  continuously batched mix, alone, and through the plain per-request path.

Sizes are those of bench.py's workloads; each test prints the case before
its first GPU step.
"""
import numpy as np
import pytest

import oracle
from jaccard_amd import api, batch, graphs, serve

pytestmark = pytest.mark.gpu

WEIGHTED_RTOL = 1e-6


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def check_weighted(got, want, what):
    """got (fp32) against the FP64 quotient: zeros exact, the rest <= 1e-6."""
    got = np.asarray(got, np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(want)), what
    ok = ~np.isnan(want)
    got, want = got[ok], want[ok]
    nz = want != 0
    assert np.all(got[~nz] == 0), f"{what}: non-zero where the intersection is empty"
    rel = np.abs(got[nz] - want[nz]) / np.abs(want[nz])
    worst = float(rel.max(initial=0.0))
    print(f"[gpu]   {what}: {got.size} values, worst rel err {worst:.3e}", flush=True)
    assert worst <= WEIGHTED_RTOL, what


def mirror_positions(dg, torch):
    """Position of (v, u) for every stored (u, v), on the device (CSR rows are
    sorted, so keys u * n + v are ascending)."""
    n = dg.n
    deg = (dg.offsets[1:] - dg.offsets[:-1]).long()
    rows = torch.repeat_interleave(torch.arange(n, device=dg.offsets.device), deg)
    cols = dg.indices.long()
    key = rows * n + cols
    mirror = torch.searchsorted(key, cols * n + rows)
    del key
    assert torch.equal(mirror.clamp(max=dg.m - 1), mirror), "mirror outside the array"
    return rows, mirror


@pytest.fixture(scope="module")
def rmat22(cuda):
    g = graphs.rmat(22, 16, seed=1, weighted=True)
    print(f"[gpu] RMAT-22 weighted n={g.n} m={g.m}", flush=True)
    dg = api.DeviceCSR.from_host(g, cuda)
    W = api.vertex_weights(dg)
    return g, dg, W


def test_rmat22_weighted_full_size(rmat22, cuda):
    import torch
    g, dg, W = rmat22
    print("[gpu] RMAT-22 weighted EC and VC", flush=True)
    ec = api.jaccard(dg, weights=W, edge_centric=True)
    vc = api.jaccard(dg, weights=W, edge_centric=False)
    assert torch.equal(ec.view(torch.int32), vc.view(torch.int32)), "weighted EC and VC differ"
    del vc
    rows, mirror = mirror_positions(dg, torch)
    assert torch.equal(ec.view(torch.int32)[mirror], ec.view(torch.int32)), "J(u,v) != J(v,u)"
    assert bool(((ec >= 0) & (ec <= 1)).all()), "J outside [0, 1]"
    del rows, mirror
    # oracle: a random sample and every edge of the five largest hubs
    rng = np.random.default_rng(22)
    deg = np.diff(g.offsets.astype(np.int64))
    hubs = np.argsort(-deg)[:5]
    hub_edges = np.concatenate([np.arange(g.offsets[h], g.offsets[h + 1]) for h in hubs])
    k = np.unique(np.concatenate([rng.integers(0, g.m, 200_000), hub_edges]))
    r = (np.searchsorted(g.offsets, k, side="right") - 1).astype(np.int32)
    c = g.indices[k].astype(np.int32)
    print(f"[gpu]   oracle on {k.size} edges (hub degrees {deg[hubs].tolist()})", flush=True)
    Wh = W.cpu().numpy().astype(np.float64)
    check_weighted(ec.cpu().numpy()[k], oracle.pairs_weighted_f64(g, Wh, r, c), "RMAT-22 weighted EC")
    print("[gpu] RMAT-22 unweighted EC on the same CSR", flush=True)
    eu = api.jaccard(dg, edge_centric=True).cpu().numpy()
    assert np.array_equal(bits(eu[k]), bits(oracle.ec_edges(g, k))), "unweighted EC vs the reference loop"


def test_chunglu20_bit_exact(cuda):
    g = graphs.chung_lu(n=1 << 20, seed=1)
    print(f"[gpu] Chung-Lu 2^20 n={g.n} m={g.m} max_deg={int(np.diff(g.offsets).max())}", flush=True)
    ref = oracle.ec_unweighted(g)
    dg = api.DeviceCSR.from_host(g, cuda)
    for ec in (True, False):
        out = api.jaccard(dg, edge_centric=ec).cpu().numpy()
        bad = np.nonzero(bits(out) != bits(ref))[0]
        assert bad.size == 0, f"ec={ec}: {bad.size} mismatches, first at {bad[:5]}"


def test_batch4096_union_equals_plain(cuda):
    import torch
    bu, bw = batch.make_batches(count=4096, seed=1)
    print(f"[gpu] batch of {bu.count + bw.count} graphs, m={bu.union.m + bw.union.m}", flush=True)
    br = batch.BatchRunner([bu, bw], cuda)
    br.step()
    pr = batch.PlainRunner([bu, bw], cuda)
    pr.step()
    torch.cuda.synchronize()
    for k, i, _, out, _ in pr.graphs:
        u = br.member_output(k, i)
        assert torch.equal(u.view(torch.int32), out.view(torch.int32)), (k, i)
    del pr
    rng = np.random.default_rng(40)
    for k, b in enumerate((bu, bw)):
        for i in rng.choice(b.count, 25, replace=False):
            g = b.member(int(i))
            got = br.member_output(k, int(i)).cpu().numpy()
            if b.union.weights is None:
                assert np.array_equal(bits(got), bits(oracle.ec_unweighted(g))), (k, int(i))
            else:
                # the vertex weights the batched step computed for this member
                Wm = br.items[k]["W"][int(b.vbase[i]):int(b.vbase[i + 1])].cpu().numpy().astype(np.float64)
                np.testing.assert_allclose(Wm, oracle.vertex_weights(g), rtol=6e-8)
                check_weighted(got, oracle.weighted_f64(g, Wm), f"batch member {int(i)}")


def test_serving64_mix_alone_plain_identical(cuda):
    import torch
    gs = serve.serving_graphs(count=64, seed=7)
    print(f"[gpu] serving: 64 graphs, m={sum(g.m for g in gs)}", flush=True)
    dev = serve.upload(gs, cuda, weighted=True)
    reqs = serve.request_stream(gs, 96, qps=0.0, seed=13)
    cont = serve.Server(gs, cuda, mode=serve.CONTINUOUS, slots=4, hbm_budget_bytes=1 << 30, device_graphs=dev)
    fifo = serve.Server(gs, cuda, mode=serve.FIFO, device_graphs=dev)
    res_c = serve.result_buffers(gs, reqs, cuda)
    st, info = cont.run(reqs, res_c)
    assert (st["status"] == 0).all()
    print(f"[gpu]   continuous: {info['batches']} batches, {info['max_in_flight']} in flight at most", flush=True)
    assert info["batches"] < len(reqs)
    res_f = serve.result_buffers(gs, reqs, cuda)
    st2, _ = fifo.run(reqs, res_f)
    assert (st2["status"] == 0).all()
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(res_c, res_f)):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (i, reqs[i].kind)
    rng = np.random.default_rng(14)
    for i in rng.choice(len(reqs), 8, replace=False):
        r1 = serve.result_buffers(gs, [reqs[i]], cuda)
        s1, _ = cont.run([reqs[i]], r1)
        assert s1["status"][0] == 0
        assert torch.equal(r1[0].view(torch.int32), res_c[i].view(torch.int32)), int(i)
    # the oracle on a few requests (pairs whole, full graphs on a sample)
    for i in rng.choice(len(reqs), 6, replace=False):
        q = reqs[i]
        g = gs[q.graph]
        Wh = dev[q.graph][1].cpu().numpy()[: g.n].astype(np.float64)
        got = res_c[i].cpu().numpy()
        if q.kind == serve.REQ_FULL:
            k = np.unique(rng.integers(0, g.m, 20_000))
            r = (np.searchsorted(g.offsets, k, side="right") - 1).astype(np.int32)
            check_weighted(got[k], oracle.pairs_weighted_f64(g, Wh, r, g.indices[k]), f"request {int(i)} (graph)")
        else:
            check_weighted(got[: q.num_pairs], oracle.pairs_weighted_f64(g, Wh, q.first, q.second),
                           f"request {int(i)} ({q.num_pairs} pairs)")
    cont.close()
    fifo.close()
