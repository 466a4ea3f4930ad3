"""Serving runtime: every request's result is the same alone, in a
continuously batched mix, and through the plain per-request path; whole-graph
results match the FP64 oracle within 1e-6."""
import numpy as np
import pytest

import oracle
from jaccard_amd import api, serve

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def setup(cuda):
    gs = serve.serving_graphs(count=8, scale_min=8, scale_max=11, edge_factor=8, seed=3)
    dev = serve.upload(gs, cuda, weighted=True)
    reqs = serve.request_stream(gs, 96, qps=0.0, p_full=0.4, kmax=4096, seed=5)
    # a few pairs outside the graph
    for q in reqs:
        if q.kind == serve.REQ_PAIRS and q.num_pairs >= 4:
            q.first[0] = -1
            q.second[1] = gs[q.graph].n
            break
    return gs, dev, reqs


def copy(res):
    return [r.cpu().numpy().copy() for r in res]


def test_mix_alone_and_plain_identical(setup, cuda):
    gs, dev, reqs = setup
    print(f"[gpu] serve: {len(reqs)} requests over {len(gs)} graphs", flush=True)
    srv = serve.Server(gs, cuda, mode=serve.CONTINUOUS, slots=3, max_batch_requests=16,
                       hbm_budget_bytes=8 << 20, device_graphs=dev)
    res = serve.result_buffers(gs, reqs, cuda)
    st, info = srv.run(reqs, res)
    assert (st["status"] == 0).all()
    assert info["batches"] > 1 and info["batches"] < len(reqs)
    mix = copy(res)
    # each request alone on the same server
    for i, q in enumerate(reqs):
        r1 = serve.result_buffers(gs, [q], cuda)
        s1, _ = srv.run([q], r1)
        assert s1["status"][0] == 0
        assert np.array_equal(bits(r1[0].cpu().numpy()), bits(mix[i])), i
    # plain per-request path
    fifo = serve.Server(gs, cuda, mode=serve.FIFO, device_graphs=dev)
    res2 = serve.result_buffers(gs, reqs, cuda)
    st2, _ = fifo.run(reqs, res2)
    assert (st2["status"] == 0).all()
    for i, r in enumerate(copy(res2)):
        a, b = bits(r), bits(mix[i])
        assert np.array_equal(a, b), (i, reqs[i].kind)
    srv.close()
    fifo.close()


def test_results_match_oracle(setup, cuda):
    gs, dev, reqs = setup
    srv = serve.Server(gs, cuda, mode=serve.CONTINUOUS, slots=2, device_graphs=dev)
    res = serve.result_buffers(gs, reqs, cuda)
    st, _ = srv.run(reqs, res)
    assert (st["status"] == 0).all()
    worst = 0.0
    for i, q in enumerate(reqs[:40]):
        g = gs[q.graph]
        W = dev[q.graph][1].cpu().numpy().astype(np.float64)
        got = res[i].cpu().numpy()
        if q.kind == serve.REQ_FULL:
            want = oracle.weighted_f64(g, W)
            got = got[:g.m]
        else:
            want = oracle.pairs_weighted_f64(g, W, q.first, q.second)
            bad = (q.first < 0) | (q.first >= g.n) | (q.second < 0) | (q.second >= g.n)
            assert np.isnan(got[bad]).all()
            got, want = got[~bad], want[~bad]
        # isolated endpoints give 0/0 = NaN, as the reference arithmetic does
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        got, want = got[ok], want[ok]
        nz = want != 0
        assert np.all(got[~nz] == 0)
        if nz.any():
            worst = max(worst, float(np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz]))))
    print(f"[gpu] serve worst rel err {worst:.3e}", flush=True)
    assert worst <= 1e-6
    srv.close()


def test_budget_gate(setup, cuda):
    gs, dev, reqs = setup
    big = max(range(len(gs)), key=lambda i: gs[i].m)
    srv = serve.Server(gs, cuda, mode=serve.CONTINUOUS, slots=2, hbm_budget_bytes=4096, device_graphs=dev)
    q = [serve.Request(big, serve.REQ_FULL, 0.0), serve.Request(0, serve.REQ_PAIRS, 0.0,
                                                                np.zeros(0, np.int32), np.zeros(0, np.int32))]
    res = serve.result_buffers(gs, q, cuda)
    st, _ = srv.run(q, res)
    assert st["status"][0] == api.ERR_OUT_OF_MEMORY
    assert st["status"][1] == 0
    srv.close()


def test_asymmetric_graph_rejected(cuda):
    """Batched passes write each pair once for both stored directions, so a
    member without its reverse edges is refused at creation (both modes)."""
    import cases
    g = cases.random_directed(n=500, m=3000, seed=9)
    print(f"[gpu] serve: asymmetric graph n={g.n} m={g.m}", flush=True)
    for mode in (serve.CONTINUOUS, serve.FIFO):
        with pytest.raises(api.JaccardError) as ex:
            serve.Server([g], cuda, weighted=False, mode=mode)
        assert ex.value.code == api.ERR_INVALID_CSR
