"""jc_jaccard_noalloc: the allocation-free entry point (fixed per-device
scratch array) against jc_jaccard_ex, bit for bit, on two calls in a row,
the second with a smaller graph; and the size limit, checked on the host
before any device work."""
import ctypes

import numpy as np
import pytest

from jaccard_amd import _native, api, graphs


def test_scratch_limit_checked_on_host():
    """A graph whose workspace exceeds the scratch is refused before any HIP
    call (so no device is needed): RMAT-21-sized counts, edge-centric
    weighted (24 bytes per vertex and more)."""
    lib = _native.load_jaccard_lib()
    n, m = 1 << 21, 1 << 24
    assert api.workspace_size(n, m, True, True) > api.NOALLOC_SCRATCH_BYTES
    fake = api.JcCsr(n, m, 0x1000, 0x2000, None)  # never dereferenced: refused first
    code = lib.jc_jaccard_noalloc(ctypes.byref(fake), 0x3000, 0x4000, 1, 0, None)
    assert code == api.ERR_SCRATCH_LIMIT
    assert b"scratch" in lib.jc_strerror(code)
    # a graph that fits is not refused for size (here: fails later, without a device)
    small = api.JcCsr(1000, 1000, 0x1000, 0x2000, None)
    assert api.workspace_size(1000, 1000, True, True) <= api.NOALLOC_SCRATCH_BYTES
    assert lib.jc_jaccard_noalloc(ctypes.byref(small), None, None, 1, 0, None) == api.ERR_INVALID_ARG


@pytest.mark.gpu
def test_noalloc_matches_ex_twice_then_smaller(cuda):
    import torch
    big = graphs.rmat(16, 16, seed=3, weighted=True)
    small = graphs.rmat(11, 8, seed=4, weighted=True)
    for g in (big, small):  # two calls in a row, the second on a smaller graph
        print(f"[gpu] noalloc n={g.n} m={g.m}", flush=True)
        dg = api.DeviceCSR.from_host(g, cuda)
        W = api.vertex_weights(dg)
        for w, ec in ((W, True), (None, True), (W, False), (None, False)):
            want = api.jaccard(dg, weights=w, edge_centric=ec)
            got = torch.full_like(want, -7.0)
            api.jaccard_noalloc(dg, got, weights=w, edge_centric=ec)
            torch.cuda.synchronize()
            assert torch.equal(got.view(torch.int32), want.view(torch.int32)), (g.n, w is not None, ec)
