"""Graphs shared by the oracle and GPU parity tests."""
import numpy as np

from jaccard_amd import graphs


def complete(n):
    a = np.ones((n, n), dtype=np.int8)
    return graphs.from_dense(a)


def star(leaves):
    src = np.zeros(leaves, np.int32)
    dst = np.arange(1, leaves + 1, dtype=np.int32)
    return graphs.from_edges(leaves + 1, src, dst)


def path(n):
    return graphs.from_edges(n, np.arange(n - 1), np.arange(1, n))


def cycle(n):
    return graphs.from_edges(n, np.arange(n), (np.arange(n) + 1) % n)


def complete_bipartite(a, b):
    s, d = np.meshgrid(np.arange(a), np.arange(a, a + b), indexing="ij")
    return graphs.from_edges(a + b, s.ravel(), d.ravel())


def two_triangles():
    # triangles (0,1,2) and (1,2,3) sharing edge (1,2)
    e = np.array([[0, 1], [0, 2], [1, 2], [1, 3], [2, 3]])
    return graphs.from_edges(4, e[:, 0], e[:, 1])


def hub_graph(n=30000, extra=8, seed=3):
    """Vertex 0 adjacent to every other vertex, plus random edges among the
    others: the hub's row spans several LDS chunks and every hub pair has a
    non-empty intersection."""
    rng = np.random.default_rng(seed)
    src = [np.zeros(n - 1, np.int64)]
    dst = [np.arange(1, n)]
    s = rng.integers(1, n, size=(n - 1) * extra // 2)
    d = rng.integers(1, n, size=s.size)
    src.append(s)
    dst.append(d)
    return graphs.from_edges(n, np.concatenate(src), np.concatenate(dst), weighted=True)


def sparse_hub_graph(n=1 << 21, k=30000, extra=8, seed=4):
    """Vertex 0 adjacent to k vertices spread over 2^21 ids, plus random edges
    among them: the hub's id span is too wide for an LDS bitmap, so its row
    runs through several hash-table chunks."""
    rng = np.random.default_rng(seed)
    nb = rng.choice(np.arange(1, n), size=k, replace=False)
    src = [np.zeros(k, np.int64), nb[rng.integers(0, k, size=k * extra // 2)]]
    dst = [nb, nb[rng.integers(0, k, size=k * extra // 2)]]
    return graphs.from_edges(n, np.concatenate(src), np.concatenate(dst), weighted=True)


def span_hubs_graph(k=3000, extra=6, seed=6):
    """Three hub rows of degree k whose id spans straddle the LDS bitmap
    limits of the task kernels (about 136 KiB weighted: 8 bytes per 32 ids
    plus 8 per key; about 139 KiB unweighted: 4 bytes per 32 ids): span
    400,000 fits both, 600,000 only the unweighted bitmap, 1,300,000 neither,
    so one graph sends rows to both task lists of either mode. Random edges
    among each hub's neighbours give the pairs common neighbours."""
    rng = np.random.default_rng(seed)
    n = 1_300_000 + 16
    src, dst = [], []
    for hub, span in ((0, 400_000), (1, 600_000), (2, 1_300_000)):
        nb = np.unique(np.linspace(3, span + 2, k).astype(np.int64))
        src.append(np.full(nb.size, hub))
        dst.append(nb)
        src.append(nb[rng.integers(0, nb.size, size=nb.size * extra // 2)])
        dst.append(nb[rng.integers(0, nb.size, size=nb.size * extra // 2)])
    return graphs.from_edges(n, np.concatenate(src), np.concatenate(dst), weighted=True)


def random_directed(n=2000, m=20000, seed=5):
    """An asymmetric CSR (no reverse edges added): exercises the fix-up pass."""
    rng = np.random.default_rng(seed)
    return graphs.from_edges(n, rng.integers(0, n, m), rng.integers(0, n, m), symmetrize=False)


def with_self_loops(g):
    """Adds (u,u) to every third non-empty row, keeping rows sorted."""
    rows = []
    off = [0]
    for u in range(g.n):
        r = g.indices[g.offsets[u]:g.offsets[u + 1]]
        if u % 3 == 0 and r.size:
            r = np.sort(np.append(r, u))
        rows.append(r)
        off.append(off[-1] + r.size)
    return graphs.CSRGraph(n=g.n, offsets=np.asarray(off, np.int32),
                           indices=np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32))


def small_suite():
    return {
        "K5": complete(5),
        "K17": complete(17),
        "K130": complete(130),
        "star200": star(200),
        "path50": path(50),
        "cycle40": cycle(40),
        "K7x9": complete_bipartite(7, 9),
        "two_triangles": two_triangles(),
        "rmat8": graphs.rmat(8, 16, seed=11),
        "rmat10": graphs.rmat(10, 16, seed=12),
        "rmat12_ef32": graphs.rmat(12, 32, seed=13),
        "chunglu_4k": graphs.chung_lu(n=4096, mean_deg=16, max_deg=1024, seed=14),
    }
