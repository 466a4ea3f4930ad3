"""Graph ingest on the device (jc_csr_from_edges) builds the same CSR as the
host builder: symmetrised, self loops dropped, repeats merged, rows sorted."""
import ctypes

import numpy as np
import pytest

from jaccard_amd import api, graphs
from jaccard_amd._native import load_graph_lib

pytestmark = pytest.mark.gpu


def dev(a, cuda):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


def raw_rmat(scale, ef, seed=1):
    lib = load_graph_lib()
    e = int(round(ef * (1 << scale)))
    s = np.empty(e, np.int32)
    d = np.empty(e, np.int32)
    got = lib.jcg_rmat_edges(scale, 0, float(ef), graphs.RMAT_A, graphs.RMAT_B, graphs.RMAT_C, seed,
                             s.ctypes.data, d.ctypes.data, e)
    assert got == e
    return s, d


@pytest.mark.parametrize("sym", [True, False])
def test_random_edges_match_host(sym, cuda):
    rng = np.random.default_rng(3)
    n, e = 5000, 60000
    s = rng.integers(0, n, e).astype(np.int32)
    d = rng.integers(0, n, e).astype(np.int32)
    s[:50] = d[:50]                      # self loops
    s[100:200] = s[300:400]              # repeated pairs
    d[100:200] = d[300:400]
    print(f"[gpu] ingest random n={n} e={e} sym={sym}", flush=True)
    g = api.csr_from_edges(n, dev(s, cuda), dev(d, cuda), symmetrize=sym)
    want = graphs.from_edges(n, s, d, symmetrize=sym)
    assert np.array_equal(g.offsets.cpu().numpy(), want.offsets)
    assert np.array_equal(g.indices.cpu().numpy(), want.indices)


def test_first_occurrence_weights(cuda):
    rng = np.random.default_rng(4)
    n, e = 300, 4000
    s = rng.integers(0, n, e).astype(np.int32)
    d = rng.integers(0, n, e).astype(np.int32)
    w = rng.random(e).astype(np.float32)
    g = api.csr_from_edges(n, dev(s, cuda), dev(d, cuda), dev(w, cuda), symmetrize=True)
    # reference: originals then reverses, self loops out, first occurrence of each (src, dst)
    ss = np.r_[s, d]
    dd = np.r_[d, s]
    ww = np.r_[w, w]
    keep = ss != dd
    key = ss[keep].astype(np.int64) * n + dd[keep]
    _, first = np.unique(key, return_index=True)
    wr = ww[keep][first]
    assert np.array_equal(g.edge_w.cpu().numpy().view(np.uint32), wr.view(np.uint32))


def test_rmat_raw_draws_give_the_generator_graph(cuda):
    s, d = raw_rmat(16, 16)
    print(f"[gpu] ingest rmat16 e={s.size}", flush=True)
    g = api.csr_from_edges(1 << 16, dev(s, cuda), dev(d, cuda))
    want = graphs.rmat(16, 16, seed=1)
    assert np.array_equal(g.offsets.cpu().numpy(), want.offsets)
    assert np.array_equal(g.indices.cpu().numpy(), want.indices)
    # and Jaccard on it matches the host-built graph bit for bit
    a = api.jaccard(g).cpu().numpy()
    b = api.jaccard(api.DeviceCSR.from_host(want, cuda)).cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_empty_and_errors(cuda):
    import torch
    e = torch.empty(0, dtype=torch.int32, device=cuda)
    g = api.csr_from_edges(10, e, e)
    assert g.m == 0 and not g.offsets.cpu().numpy().any()
    bad = dev(np.array([0, 10], np.int32), cuda)
    ok = dev(np.array([1, 2], np.int32), cuda)
    with pytest.raises(api.JaccardError) as ex:
        api.csr_from_edges(10, bad, ok)
    assert ex.value.code == api.ERR_INVALID_ARG
