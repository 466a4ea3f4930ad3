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
