"""Hash-table kinds of the workgroup task kernel.

Graphs with fewer than 2^24 vertices use quotient tables (16-bit keys from
a bijective 24-bit hash, dense weights); larger graphs, or the tuning switch
``hash_table=1``, use 32-bit tables. A task whose quotient table cannot be
built (a key finds no slot within the largest displacement) is redone with
the 32-bit table. Every variant must give the oracle's bits (unweighted) and
the same bits as the others (weighted, where the FP64 oracle bounds them).
"""
import numpy as np
import pytest

import oracle
from jaccard_amd import api, graphs

import cases

pytestmark = pytest.mark.gpu

WEIGHTED_RTOL = 1e-6
HASH_MUL = 0x9E3779  # the engine's 24-bit Fibonacci multiplier
VARIANTS = {"default": None, "table32": {"hash_table": 1}, "forced_redo": {"quot_max_disp": 1}}


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def colliding_hub_graph(seed=8):
    """n = 2^24 - 1 (the largest quotient-table graph). Vertex 0's row holds
    40 ids whose 24-bit hashes share their top 10 bits (one home bucket of a
    1024-bucket table, which with its two displacement buckets holds 21
    keys) plus 160 random ids: its table cannot be built and its task is
    redone with the 32-bit table. Random edges among the row's ids give the
    pairs common neighbours."""
    rng = np.random.default_rng(seed)
    n = (1 << 24) - 1
    inv = pow(HASH_MUL, -1, 1 << 24)
    home = 613
    hs = (home << 14) + rng.choice(1 << 14, 400, replace=False)
    coll = np.array([(inv * int(h)) % (1 << 24) for h in hs], np.int64)
    coll = coll[(coll > 0) & (coll < n)][:40]
    assert coll.size == 40
    rest = rng.choice(np.arange(1, n), 160, replace=False)
    nb = np.unique(np.concatenate([coll, rest]))
    src = [np.zeros(nb.size, np.int64), nb[rng.integers(0, nb.size, 1200)]]
    dst = [nb, nb[rng.integers(0, nb.size, 1200)]]
    return graphs.from_edges(n, np.concatenate(src), np.concatenate(dst), weighted=True)


def wide_id_graph(seed=9):
    """n = 2^24 + 5: the 32-bit tables by size. A hub of degree 9000 over the
    whole id range (three weighted chunks), with ids at and above 2^24, and
    random edges among its neighbours."""
    rng = np.random.default_rng(seed)
    n = (1 << 24) + 5
    nb = np.unique(np.concatenate([rng.choice(np.arange(1, n), 8990, replace=False),
                                   np.arange(n - 10, n)]))
    src = [np.zeros(nb.size, np.int64), nb[rng.integers(0, nb.size, 27000)]]
    dst = [nb, nb[rng.integers(0, nb.size, 27000)]]
    return graphs.from_edges(n, np.concatenate(src), np.concatenate(dst), weighted=True)


GRAPHS = {
    "rmat14": lambda: graphs.rmat(14, 16, seed=21, weighted=True),
    "hub_hash_chunks": cases.sparse_hub_graph,
    "colliding_hub": colliding_hub_graph,
    "wide_ids": wide_id_graph,
}


@pytest.mark.parametrize("name", list(GRAPHS))
def test_table_kinds_agree(name, cuda):
    g = GRAPHS[name]()
    print(f"[gpu] table kinds {name} n={g.n} m={g.m} max_deg={int(np.diff(g.offsets).max())}", flush=True)
    dg = api.DeviceCSR.from_host(g, cuda)
    ref = oracle.ec_unweighted(g)
    W = api.vertex_weights(dg)
    refw = oracle.weighted_f64(g, W.cpu().numpy().astype(np.float64))
    first = None
    for vname, tun in VARIANTS.items():
        for ec in (True, False):
            print(f"[gpu]   {vname} ec={ec}", flush=True)
            out = api.jaccard(dg, edge_centric=ec, tuning=tun).cpu().numpy()
            bad = np.nonzero(bits(out) != bits(ref))[0]
            assert bad.size == 0, f"{vname} ec={ec}: {bad.size} mismatches, first at {bad[:5]}"
            ow = api.jaccard(dg, weights=W, edge_centric=ec, tuning=tun).cpu().numpy()
            ok = refw != 0
            rel = np.abs(ow[ok].astype(np.float64) - refw[ok]) / np.abs(refw[ok])
            assert rel.max(initial=0.0) <= WEIGHTED_RTOL, f"{vname} ec={ec}"
            assert np.all(ow[~ok] == 0)
            if first is None:
                first = ow
            assert np.array_equal(bits(ow), bits(first)), f"{vname} ec={ec}: weighted bits differ"
