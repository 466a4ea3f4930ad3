"""Host-side CSR graphs: deterministic generators and the int32 CSR container.

The CSR layout is the one the reference's kernels consume
(``GraphCSRView<int32_t,int32_t,float>``, standalone_csr.hpp:44-105):
``offsets[n+1]`` int32, ``indices[m]`` int32 (sorted inside each row),
optional ``weights[m]`` float32. Graphs are symmetrised, deduplicated and free
of self loops, as main.cpp:98-120 and readMtxToCSR.cpp:152-208 prepare them.

Generation runs in the native host library ``libjcgraph.so`` (C++/OpenMP,
counter-based randomness), so every graph is a pure function of its
parameters.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Optional

import numpy as np

from ._native import load_graph_lib

# RMAT quadrant probabilities used by every workload here (d = 0.05).
RMAT_A, RMAT_B, RMAT_C = 0.57, 0.19, 0.19
DEFAULT_WSEED_XOR = 0x57454947


@dataclasses.dataclass
class CSRGraph:
    """An int32 CSR graph resident in host memory."""

    n: int
    offsets: np.ndarray  # int32[n+1]
    indices: np.ndarray  # int32[m]
    weights: Optional[np.ndarray] = None  # float32[m] or None
    name: str = ""

    @property
    def m(self) -> int:
        return int(self.indices.shape[0])

    @property
    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets.astype(np.int64))

    def stats(self) -> dict:
        """Exact int64 work statistics (sum of min/sum degrees over stored edges)."""
        lib = load_graph_lib()
        out = (ctypes.c_int64 * 4)()
        lib.jcg_stats(self.n, _ptr(self.offsets), _ptr(self.indices), out)
        return {"sum_min_deg": int(out[0]), "sum_deg_pair": int(out[1]),
                "max_deg": int(out[2]), "isolated": int(out[3])}

    def validate(self) -> None:
        """Raise ValueError unless this is a well-formed int32 CSR."""
        off, ind = self.offsets, self.indices
        if off.dtype != np.int32 or ind.dtype != np.int32:
            raise ValueError("offsets and indices must be int32")
        if off.shape != (self.n + 1,) or off[0] != 0 or off[-1] != ind.shape[0]:
            raise ValueError("offsets must have n+1 entries, start at 0 and end at m")
        if np.any(np.diff(off.astype(np.int64)) < 0):
            raise ValueError("offsets must be non-decreasing")
        if ind.size and (ind.min() < 0 or ind.max() >= self.n):
            raise ValueError("column index out of range")
        if self.weights is not None and (self.weights.dtype != np.float32
                                         or self.weights.shape != ind.shape):
            raise ValueError("weights must be float32[m]")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _take(handle, name: str) -> CSRGraph:
    lib = load_graph_lib()
    if not handle:
        raise ValueError(f"graph generation failed for {name} (bad parameters or m > 2^31-1)")
    try:
        n = lib.jcg_n(handle)
        m = lib.jcg_m(handle)
        off = np.ctypeslib.as_array(lib.jcg_offsets(handle), shape=(n + 1,)).copy()
        ind = (np.ctypeslib.as_array(lib.jcg_indices(handle), shape=(m,)).copy()
               if m else np.zeros(0, np.int32))
        w = None
        if lib.jcg_weighted(handle):
            w = (np.ctypeslib.as_array(lib.jcg_weights(handle), shape=(m,)).copy()
                 if m else np.zeros(0, np.float32))
        return CSRGraph(n=n, offsets=off, indices=ind, weights=w, name=name)
    finally:
        lib.jcg_free(handle)


def rmat(scale: int, edge_factor: float = 16, seed: int = 1, weighted: bool = False,
         n: Optional[int] = None, a: float = RMAT_A, b: float = RMAT_B, c: float = RMAT_C,
         wseed: Optional[int] = None) -> CSRGraph:
    """RMAT graph: round(edge_factor*n) draws, symmetrised, deduped, sorted.

    ``n`` defaults to 2**scale; a smaller ``n`` keeps only draws below it.
    """
    lib = load_graph_lib()
    nn = 0 if n is None else int(n)
    ws = (seed ^ DEFAULT_WSEED_XOR) if wseed is None else wseed
    h = lib.jcg_rmat(int(scale), nn, float(edge_factor), a, b, c, ctypes.c_uint64(seed),
                     1 if weighted else 0, ctypes.c_uint64(ws))
    tag = f"rmat{scale}" + (f"_n{n}" if n else "") + f"_ef{edge_factor}_s{seed}" + ("_w" if weighted else "")
    return _take(h, tag)


def rmat_n(n: int, edge_factor: float, seed: int, weighted: bool = False) -> CSRGraph:
    """RMAT graph on exactly ``n`` vertices (scale = ceil(log2 n))."""
    scale = max(1, math.ceil(math.log2(n)))
    return rmat(scale, edge_factor, seed, weighted, n=n)


def chung_lu(n: int = 1 << 20, gamma: float = 2.1, mean_deg: float = 32.0,
             max_deg: float = float(1 << 17), seed: int = 1, weighted: bool = False) -> CSRGraph:
    """Chung-Lu power-law graph calibrated to the given mean stored degree."""
    lib = load_graph_lib()
    h = lib.jcg_chung_lu(int(n), float(gamma), float(mean_deg), float(max_deg),
                         ctypes.c_uint64(seed), 1 if weighted else 0,
                         ctypes.c_uint64(seed ^ DEFAULT_WSEED_XOR))
    return _take(h, f"chunglu_n{n}_g{gamma}_d{mean_deg}_s{seed}" + ("_w" if weighted else ""))


def from_edges(n: int, src, dst, symmetrize: bool = True, weighted: bool = False,
               wseed: int = DEFAULT_WSEED_XOR) -> CSRGraph:
    """CSR from a directed edge list (dedupe, drop self loops, optional symmetrise)."""
    lib = load_graph_lib()
    s = np.ascontiguousarray(src, dtype=np.int32)
    d = np.ascontiguousarray(dst, dtype=np.int32)
    if s.shape != d.shape:
        raise ValueError("src and dst must have the same length")
    h = lib.jcg_from_edges(int(n), int(s.size), _ptr(s), _ptr(d), 1 if symmetrize else 0,
                           1 if weighted else 0, ctypes.c_uint64(wseed))
    return _take(h, f"edges_n{n}")


def from_dense(adj: np.ndarray, weights: Optional[np.ndarray] = None) -> CSRGraph:
    """CSR of a dense 0/1 adjacency matrix taken as given (no symmetrisation).

    ``weights`` (same shape, float32) supplies per-entry edge weights.
    """
    a = np.asarray(adj) != 0
    n = a.shape[0]
    np.fill_diagonal(a, False)
    rows, cols = np.nonzero(a)
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, rows + 1, 1)
    off = np.cumsum(off).astype(np.int32)
    w = None if weights is None else np.asarray(weights, np.float32)[rows, cols].copy()
    return CSRGraph(n=n, offsets=off, indices=cols.astype(np.int32), weights=w, name="dense")


def pair_weight(wseed: int, u: int, v: int) -> float:
    """The symmetric per-pair weight the generators attach to edge (u, v)."""
    return float(load_graph_lib().jcg_pair_weight(ctypes.c_uint64(wseed), int(u), int(v)))
