"""jc_degree_stats against a host histogram, and jc_select_config's choice
against default and forced launch configurations: same bits."""
import numpy as np
import pytest

from jaccard_amd import api, graphs


def host_bitmap_rows(g, table_bytes):
    """Long rows that get a bitmap table, by the rule jaccard_amd.h states."""
    off = g.offsets.astype(np.int64)
    deg = np.diff(off)
    longr = np.nonzero(deg > 128)[0]
    lo = g.indices[off[longr]].astype(np.int64)
    hi = g.indices[off[longr + 1] - 1].astype(np.int64)
    words = (hi - lo) // 32 + 2
    d = deg[longr]
    return (int((4 * words <= table_bytes[0]).sum()),
            int(((8 * words + 8 * (d + 1) <= table_bytes[1]) & (8 * words <= 32 * 1024)).sum()))


def host_stats(g):
    deg = np.diff(g.offsets.astype(np.int64))
    rows, ent = np.zeros(32, np.int64), np.zeros(32, np.int64)
    nz = deg[deg > 0]
    b = np.floor(np.log2(nz)).astype(np.int64)
    np.add.at(rows, b, 1)
    np.add.at(ent, b, nz)
    return rows, ent, int((deg == 0).sum()), int((deg > 128).sum()), int(deg.max(initial=0))


CASES = {
    "rmat14": lambda: graphs.rmat(14, 16, seed=2),
    "rmat13_isolated": lambda: graphs.rmat(13, 2, seed=3, n=5000),
    "chunglu_2e15": lambda: graphs.chung_lu(1 << 15, seed=4),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_degree_stats_match_host(name, cuda):
    g = CASES[name]()
    print(f"[gpu] degree stats {name} n={g.n} m={g.m}", flush=True)
    dg = api.DeviceCSR.from_host(g, cuda)
    st = api.degree_stats(dg)
    rows, ent, iso, longr, dmax = host_stats(g)
    assert list(st.rows) == rows.tolist()
    assert list(st.entries) == ent.tolist()
    assert st.isolated == iso and st.long_rows == longr and st.max_degree == dmax
    assert list(st.long_bitmap_rows) == list(host_bitmap_rows(g, st.bitmap_table_bytes))
    # at most 2^16 ids: every long row's unweighted bitmap fits
    assert st.long_bitmap_rows[0] == longr


@pytest.mark.gpu
def test_degree_stats_wide_span(cuda):
    """Long rows spanning 2^22 ids do not fit a bitmap: counted as long only."""
    n = 1 << 22
    src = np.repeat(np.arange(4, dtype=np.int32), 200)
    dst = (np.tile(np.arange(200, dtype=np.int64), 4) * (n // 200) + 7).astype(np.int32)
    g = graphs.from_edges(n, src, dst)
    dg = api.DeviceCSR.from_host(g, cuda)
    st = api.degree_stats(dg)
    assert st.long_rows == 4
    assert st.long_bitmap_rows[0] == 0 and st.long_bitmap_rows[1] == 0
    assert list(st.long_bitmap_rows) == list(host_bitmap_rows(g, st.bitmap_table_bytes))


@pytest.mark.gpu
@pytest.mark.parametrize("weighted", [False, True])
def test_selected_config_same_bits(weighted, cuda):
    import torch
    for g in (graphs.rmat(14, 16, seed=5, weighted=weighted), graphs.chung_lu(1 << 14, seed=6, weighted=weighted)):
        print(f"[gpu] select config n={g.n} m={g.m} weighted={weighted}", flush=True)
        dg = api.DeviceCSR.from_host(g, cuda)
        W = api.vertex_weights(dg) if weighted else None
        ec, tuning = api.select_config(dg, weighted)
        assert ec is True
        chosen = api.jaccard(dg, weights=W, edge_centric=ec, tuning=tuning)
        ref = api.jaccard(dg, weights=W, edge_centric=True)
        assert torch.equal(chosen.view(torch.int32), ref.view(torch.int32))
        for forced in ({"small_max": 64}, {"hash_table": 1}, {"bitmap_blocks_per_cu": 2},
                       {"small_max": 32, "small_blocks_per_cu": 8}):
            out = api.jaccard(dg, weights=W, edge_centric=False, tuning=forced)
            assert torch.equal(out.view(torch.int32), ref.view(torch.int32)), forced
