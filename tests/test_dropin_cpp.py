"""The C++ drop-in header (include/sygraph_jaccard.hpp) in a real host program.

tests/cpp/dropin_check.cpp is written against the header the way the
reference's driver calls its algorithm API (standalone_algorithms.hpp:67-69,
:86-87, :109-112, :131-133; main.cpp:211-275), and is built next to it by
``__graft_entry__.build()`` (tests/cpp/Makefile). On the CPU: the header and
program compile, and every overload reports an argument error as
``sygraph::jaccard_error`` without touching a device. On the GPU: every
overload runs on a committed golden fixture and returns its expected values.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "dropin_check")
GOLDEN = os.path.join(ROOT, "tests", "golden")
WEIGHTED_RTOL = 1e-6


def _binary():
    if not os.path.exists(BIN):
        raise RuntimeError(f"{BIN} is not built: run __graft_entry__.build() (make -C tests/cpp)")
    return BIN


def test_header_program_compiles():
    r = subprocess.run(["make", "-s", "-C", CPP, "syntax"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_every_overload_reports_argument_errors():
    r = subprocess.run([_binary(), "--errors"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "7 of 7" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("fixture", ["rmat9", "chunglu_1k", "directed_200"])
def test_every_overload_on_golden_fixture(fixture):
    z = np.load(os.path.join(GOLDEN, fixture + ".npz"))
    n, off, ind = int(z["n"]), z["offsets"].astype(np.int32), z["indices"].astype(np.int32)
    m, k = ind.size, z["first"].size
    print(f"[gpu] drop-in program on {fixture} n={n} m={m} k={k}", flush=True)
    ew = z["edge_w"].astype(np.float32) if "edge_w" in z.files else np.ones(m, np.float32)
    vw = z["vertex_w"].astype(np.float32)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([n, m, k], np.int32).tofile(f)
            for a in (off, ind, ew, vw, z["first"].astype(np.int32), z["second"].astype(np.int32)):
                a.tofile(f)
        r = subprocess.run([_binary(), fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        out = np.fromfile(fout, np.float32)
    assert out.size == n + 4 * m + 2 * k
    W, rest = out[:n], out[n:]
    ecw, vcw, ecu, vcu = (rest[i * m:(i + 1) * m] for i in range(4))
    lw, lu = rest[4 * m:4 * m + k], rest[4 * m + k:]

    def bits(a):
        return np.ascontiguousarray(a, np.float32).view(np.uint32)

    if "vertex_w_expected" in z.files:
        np.testing.assert_allclose(W, z["vertex_w_expected"], rtol=6e-8)
    want_u = z["expected_unweighted"].astype(np.float32)
    assert np.array_equal(bits(ecu), bits(want_u)), "jaccard<true>(graph, result)"
    assert np.array_equal(bits(vcu), bits(want_u)), "jaccard<false>(graph, result)"
    want_w = z["expected_weighted"]
    for name, got in (("jaccard<true>(graph, weights, result)", ecw), ("jaccard<false>(graph, weights, result)", vcw)):
        fin_ = np.isfinite(want_w)
        assert np.array_equal(np.isnan(got), np.isnan(want_w)), name
        rel = np.abs(got[fin_] - want_w[fin_]) / np.maximum(np.abs(want_w[fin_]), 1e-300)
        assert rel.max(initial=0.0) <= WEIGHTED_RTOL, name
    assert np.array_equal(bits(ecw), bits(vcw)), "weighted EC and VC differ"
    pu = z["expected_pairs_unweighted"].astype(np.float32)
    assert np.array_equal(bits(lu), bits(pu)), "jaccard_list(graph, n, first, second, result)"
    pw = z["expected_pairs_weighted"]
    assert np.array_equal(np.isnan(lw), np.isnan(pw)), "jaccard_list weighted NaN positions"
    f = np.isfinite(pw)
    rel = np.abs(lw[f] - pw[f]) / np.maximum(np.abs(pw[f]), 1e-300)
    assert rel.max(initial=0.0) <= WEIGHTED_RTOL, "jaccard_list(graph, weights, ...)"
