"""Batched form: one union launch (and its captured HIP graph) gives the same
bits as one plain call per graph."""
import numpy as np
import pytest

import oracle
from jaccard_amd import api, batch

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def small_batches():
    return batch.make_batches(count=96, seed=3, nmin=1 << 5, nmax=1 << 11, efmin=4, efmax=24)


def test_union_matches_plain_and_oracle(small_batches, cuda):
    import torch
    bu, bw = small_batches
    print(f"[gpu] batch members {bu.count}+{bw.count} m={bu.union.m}+{bw.union.m}", flush=True)
    br = batch.BatchRunner([bu, bw], cuda)
    br.step()
    pr = batch.PlainRunner([bu, bw], cuda)
    pr.step()
    torch.cuda.synchronize()
    for k, i, dg, out, W in pr.graphs:
        u = br.member_output(k, i).cpu().numpy()
        assert np.array_equal(bits(u), bits(out.cpu().numpy())), (k, i)
    # a few members against the oracle too
    for i in (0, bu.count // 2, bu.count - 1):
        g = bu.member(i)
        assert np.array_equal(bits(br.member_output(0, i).cpu().numpy()), bits(oracle.ec_unweighted(g)))


def test_tuning_does_not_change_results(small_batches, cuda):
    import torch
    bu, bw = small_batches
    br = batch.BatchRunner([bu, bw], cuda)
    ref = None
    for cand in batch.BatchRunner.TUNING_SPACE:
        br.tuning = [cand, cand]
        br.step()
        torch.cuda.synchronize()
        cur = [it["out"].cpu().numpy().copy() for it in br.items]
        if ref is None:
            ref = cur
        for a, b in zip(cur, ref):
            assert np.array_equal(bits(a), bits(b)), cand


def test_graph_replays_match_eager(small_batches, cuda):
    import torch
    bu, bw = small_batches
    br = batch.BatchRunner([bu, bw], cuda)
    br.autotune(reps=1)
    br.step()
    torch.cuda.synchronize()
    eager = [it["out"].cpu().numpy().copy() for it in br.items]
    br.capture()
    for rep in range(3):
        for it in br.items:
            it["out"].fill_(-1.0)        # eager work between replays
        torch.cuda.synchronize()
        br.replay()
        torch.cuda.synchronize()
        for it, e in zip(br.items, eager):
            assert np.array_equal(bits(it["out"].cpu().numpy()), bits(e)), rep
