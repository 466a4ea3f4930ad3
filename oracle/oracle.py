"""ctypes bindings of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker (or the timed host baseline); never by the product path. The C
restatement and the reference lines each function follows are documented in
oracle/jaccard_oracle.c.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        vp, i32, i64, I = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int
        L.or_ec_unweighted.argtypes = [i32, i64, vp, vp, vp, I]
        L.or_ec_edges.argtypes = [i32, vp, vp, i64, vp, vp, I]
        L.or_vc_unweighted.argtypes = [i32, i64, vp, vp, vp, I]
        L.or_vertex_weights.argtypes = [i32, vp, vp, vp, I]
        L.or_weighted_f64.argtypes = [i32, i64, vp, vp, vp, vp, I]
        L.or_pairs_unweighted.argtypes = [i32, vp, vp, i64, vp, vp, vp, I]
        L.or_pairs_weighted_f64.argtypes = [i32, vp, vp, vp, i64, vp, vp, vp, I]
        L.or_nonzero_count.argtypes = [i64, vp, ctypes.c_float]
        L.or_nonzero_count.restype = i64
        L.or_max_threads.restype = I
        for f in (L.or_ec_unweighted, L.or_ec_edges, L.or_vc_unweighted, L.or_vertex_weights,
                  L.or_weighted_f64, L.or_pairs_unweighted, L.or_pairs_weighted_f64):
            f.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def ec_unweighted(g, threads: int = 0) -> np.ndarray:
    out = np.empty(g.m, np.float32)
    lib().or_ec_unweighted(g.n, g.m, _p(_c(g.offsets, np.int32)), _p(_c(g.indices, np.int32)), _p(out), threads)
    return out


def ec_edges(g, edges, threads: int = 0) -> np.ndarray:
    e = _c(edges, np.int64)
    out = np.empty(e.size, np.float32)
    lib().or_ec_edges(g.n, _p(_c(g.offsets, np.int32)), _p(_c(g.indices, np.int32)), e.size, _p(e), _p(out), threads)
    return out


def vc_unweighted(g, threads: int = 0) -> np.ndarray:
    out = np.empty(g.m, np.float32)
    lib().or_vc_unweighted(g.n, g.m, _p(_c(g.offsets, np.int32)), _p(_c(g.indices, np.int32)), _p(out), threads)
    return out


def vertex_weights(g, threads: int = 0) -> np.ndarray:
    W = np.empty(g.n, np.float64)
    lib().or_vertex_weights(g.n, _p(_c(g.offsets, np.int32)), _p(_c(g.weights, np.float32)), _p(W), threads)
    return W


def weighted_f64(g, v, threads: int = 0) -> np.ndarray:
    out = np.empty(g.m, np.float64)
    lib().or_weighted_f64(g.n, g.m, _p(_c(g.offsets, np.int32)), _p(_c(g.indices, np.int32)),
                          _p(_c(v, np.float64)), _p(out), threads)
    return out


def pairs_unweighted(g, first, second, threads: int = 0) -> np.ndarray:
    f, s = _c(first, np.int32), _c(second, np.int32)
    out = np.empty(f.size, np.float32)
    lib().or_pairs_unweighted(g.n, _p(_c(g.offsets, np.int32)), _p(_c(g.indices, np.int32)), f.size, _p(f), _p(s),
                              _p(out), threads)
    return out


def pairs_weighted_f64(g, v, first, second, threads: int = 0) -> np.ndarray:
    f, s = _c(first, np.int32), _c(second, np.int32)
    out = np.empty(f.size, np.float64)
    lib().or_pairs_weighted_f64(g.n, _p(_c(g.offsets, np.int32)), _p(_c(g.indices, np.int32)), _p(_c(v, np.float64)),
                                f.size, _p(f), _p(s), _p(out), threads)
    return out


def nonzero_count(out, thresh: float = 1e-5) -> int:
    a = _c(out, np.float32)
    return int(lib().or_nonzero_count(a.size, _p(a), thresh))


def max_threads() -> int:
    return int(lib().or_max_threads())
