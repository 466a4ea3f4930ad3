"""Host side of the serving workload: request stream shape and latency summary (CPU)."""
import numpy as np

from jaccard_amd import graphs, serve


def small_graphs():
    return [graphs.rmat(s, 8, seed=s) for s in (7, 8, 9)]


def test_request_stream_shape():
    gs = small_graphs()
    reqs = serve.request_stream(gs, 400, qps=1000.0, p_full=0.3, kmax=1024, seed=1)
    arr = np.array([q.arrival for q in reqs])
    assert np.all(np.diff(arr) >= 0)
    # Poisson at 1000/s: mean gap 1 ms
    assert 0.7e-3 < np.diff(arr).mean() < 1.3e-3
    kinds = np.array([q.kind for q in reqs])
    assert 0.2 < (kinds == serve.REQ_FULL).mean() < 0.4
    ks = np.array([q.num_pairs for q in reqs if q.kind == serve.REQ_PAIRS])
    assert ks.min() >= 1 and ks.max() <= 1024
    assert np.median(ks) < 200  # log-uniform, not uniform
    for q in reqs:
        if q.kind != serve.REQ_PAIRS:
            continue
        g = gs[q.graph]
        assert q.first.dtype == np.int32 and q.first.size == q.second.size
        assert (q.first >= 0).all() and (q.first < g.n).all() and (q.second < g.n).all()
        # the first half of the draws are stored edges
        rows = np.repeat(np.arange(g.n), np.diff(g.offsets))
        edges = set(zip(rows.tolist(), g.indices.tolist()))
        hits = sum((int(a), int(b)) in edges for a, b in zip(q.first, q.second))
        assert hits >= q.num_pairs // 2


def test_request_stream_deterministic_and_burst():
    gs = small_graphs()
    a = serve.request_stream(gs, 50, qps=0.0, seed=4)
    b = serve.request_stream(gs, 50, qps=0.0, seed=4)
    assert all(x.arrival == 0.0 for x in a)
    for x, y in zip(a, b):
        assert x.graph == y.graph and x.kind == y.kind
        if x.kind == serve.REQ_PAIRS:
            assert np.array_equal(x.first, y.first) and np.array_equal(x.second, y.second)


def test_latency_summary():
    reqs = [serve.Request(0, serve.REQ_FULL, 0.0), serve.Request(0, serve.REQ_PAIRS, 1.0, np.zeros(1, np.int32),
                                                                   np.zeros(1, np.int32))]
    st = np.zeros(2, dtype=[("admitted", "f8"), ("completed", "f8"), ("batch", "i4"), ("status", "i4")])
    st["completed"] = [0.5, 1.25]
    d = serve.latency_summary(reqs, st)
    assert d["wall_s"] == 1.25 and d["full_p50_ms"] == 500.0 and d["pairs_p50_ms"] == 250.0
