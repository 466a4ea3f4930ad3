"""Graph files of the reference workflow: MatrixMarket coordinate text (.mtx)
and the binary CSR format version 2 (.csr), read and written by the native
host library (host/graphio.cpp; formats of readMtxToCSR.cpp:48-109, 285-395).

``load`` prepares a graph the way the reference driver does before the
Jaccard call (main.cpp:71-120): MTX weights are read as 1.0 unless
``keep_mtx_weights``; undirected inputs without stored reverse edges get
them; self loops and repeated entries are dropped for MTX inputs.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Optional

import numpy as np

from . import graphs
from ._native import load_graph_lib

# flag bits of the binary CSR header (bit 0 upward)
WEIGHTED, ZERO_INDEXED, DIRECTED, HAS_REVERSE_EDGES = 1, 2, 4, 8
VERTEX_T64, EDGE_T64, WEIGHT_T64 = 16, 32, 64


class FileFormatError(RuntimeError):
    def __init__(self, status: int, path: str):
        super().__init__(f"cannot read {path!r} (status {status})")
        self.status = status


@dataclasses.dataclass
class LoadedGraph:
    graph: graphs.CSRGraph
    flags: int

    @property
    def weighted(self) -> bool:
        return bool(self.flags & WEIGHTED)

    @property
    def directed(self) -> bool:
        return bool(self.flags & DIRECTED)


def load(path: str, add_reverse: bool = True, keep_mtx_weights: bool = False) -> LoadedGraph:
    lib = load_graph_lib()
    st = ctypes.c_int(0)
    h = lib.jcio_load(os.fsencode(path), 1 if add_reverse else 0, 0 if keep_mtx_weights else 1, ctypes.byref(st))
    if not h:
        raise FileFormatError(st.value, path)
    try:
        n, m, flags = int(lib.jcio_n(h)), int(lib.jcio_m(h)), int(lib.jcio_flags(h))
        off = np.ctypeslib.as_array(lib.jcio_offsets(h), shape=(n + 1,)).copy()
        ind = np.ctypeslib.as_array(lib.jcio_indices(h), shape=(max(m, 1),))[:m].copy()
        wp = lib.jcio_weights(h)
        w = np.ctypeslib.as_array(wp, shape=(max(m, 1),))[:m].copy() if wp and m else None
    finally:
        lib.jcio_free(h)
    g = graphs.CSRGraph(n=n, offsets=off.astype(np.int32), indices=ind.astype(np.int32),
                        weights=w if (flags & WEIGHTED) else None, name=os.path.basename(path))
    return LoadedGraph(g, flags)


def save(path: str, g: graphs.CSRGraph, values: Optional[np.ndarray] = None, directed: bool = False,
         remove_reverse: bool = False, flags: int = ZERO_INDEXED | HAS_REVERSE_EDGES) -> None:
    """Writes g as .csr or .mtx; ``values`` (one per stored edge, e.g. Jaccard
    results) become the edge weights (default: g.weights). ``remove_reverse``
    keeps one direction of each undirected pair (src > dst)."""
    lib = load_graph_lib()
    w = values if values is not None else g.weights
    off = np.ascontiguousarray(g.offsets, np.int32)
    ind = np.ascontiguousarray(g.indices, np.int32)
    wa = None if w is None else np.ascontiguousarray(w, np.float32)
    fl = flags | (DIRECTED if directed else 0)
    rc = lib.jcio_save(os.fsencode(path), g.n, g.m, off.ctypes.data, ind.ctypes.data,
                       None if wa is None else wa.ctypes.data, fl, 1 if remove_reverse else 0)
    if rc != 0:
        raise FileFormatError(rc, path)
