"""bench.py contract checks that need no GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_multi_gpu_request_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_algorithmic_bytes_formula():
    import bench
    from jaccard_amd import graphs
    g = graphs.rmat(8, 8, seed=1)
    st = g.stats()
    b = bench.algorithmic_bytes(g, st, weighted=False)
    assert b == 4 * st["sum_min_deg"] + 4 * (g.n + 1) + 8 * g.m
    assert bench.algorithmic_bytes(g, st, weighted=True) == b + 4 * g.m + 8 * g.n


def test_oracle_check_reports_errors():
    """The bench's sampled oracle check: exact fp32 roundings of the FP64
    values pass, a planted error in one sampled position is reported."""
    import numpy as np
    import torch
    import bench
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from jaccard_amd import graphs
    g = graphs.rmat(9, 8, seed=2, weighted=True)
    W = oracle.vertex_weights(g).astype(np.float32)
    ref = oracle.weighted_f64(g, W.astype(np.float64))
    out = torch.from_numpy(ref.astype(np.float32))
    ok = bench.oracle_check(g, out, torch.from_numpy(W), True, n_sample=g.m)
    assert ok["sample_edges"] == g.m and ok["max_rel_err"] <= 2 ** -24 and ok["zeros_exact"]
    bad = out.clone()
    i = int(np.flatnonzero(ref > 0)[0])
    bad[i] = bad[i] * (1 + 1e-5)
    assert bench.oracle_check(g, bad, torch.from_numpy(W), True, n_sample=g.m)["max_rel_err"] > 1e-6
    u = torch.from_numpy(oracle.ec_unweighted(g))
    assert bench.oracle_check(g, u, None, False, n_sample=g.m)["bit_exact"]
    u2 = u.clone()
    u2[i] = float(np.nextafter(np.float32(u2[i].item()), np.float32(2)))
    assert not bench.oracle_check(g, u2, None, False, n_sample=g.m)["bit_exact"]
