"""Loading of the package's native libraries (built in-tree, never a fallback).

* ``libjcgraph.so``   host C++/OpenMP: generators, CSR builder, statistics
* ``libjaccard_amd.so`` HIP for gfx950: the Jaccard kernels behind the C ABI
  declared in ``include/jaccard_amd.h``

A missing library raises ``NativeLibraryMissing``: there is no Python or CPU
substitute for the device path.
"""
from __future__ import annotations

import ctypes
import os
import threading

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GRAPH_LIB = os.path.join(PKG_DIR, "libjcgraph.so")
JACCARD_LIB = os.path.join(PKG_DIR, "libjaccard_amd.so")

_lock = threading.Lock()
_graph = None
_jaccard = None


class NativeLibraryMissing(RuntimeError):
    pass


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C sycl-jaccard_amd`")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def load_graph_lib() -> ctypes.CDLL:
    global _graph
    with _lock:
        if _graph is None:
            lib = _load(GRAPH_LIB)
            vp, i32, i64, u64, f64 = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                      ctypes.c_uint64, ctypes.c_double)
            lib.jcg_rmat.restype = vp
            lib.jcg_rmat.argtypes = [ctypes.c_int, i32, f64, f64, f64, f64, u64, ctypes.c_int, u64]
            lib.jcg_chung_lu.restype = vp
            lib.jcg_chung_lu.argtypes = [i32, f64, f64, f64, u64, ctypes.c_int, u64]
            lib.jcg_from_edges.restype = vp
            lib.jcg_from_edges.argtypes = [i32, i64, vp, vp, ctypes.c_int, ctypes.c_int, u64]
            lib.jcg_n.restype = i32
            lib.jcg_n.argtypes = [vp]
            lib.jcg_m.restype = i64
            lib.jcg_m.argtypes = [vp]
            lib.jcg_weighted.restype = ctypes.c_int
            lib.jcg_weighted.argtypes = [vp]
            lib.jcg_offsets.restype = ctypes.POINTER(i32)
            lib.jcg_offsets.argtypes = [vp]
            lib.jcg_indices.restype = ctypes.POINTER(i32)
            lib.jcg_indices.argtypes = [vp]
            lib.jcg_weights.restype = ctypes.POINTER(ctypes.c_float)
            lib.jcg_weights.argtypes = [vp]
            lib.jcg_free.restype = None
            lib.jcg_free.argtypes = [vp]
            lib.jcg_pair_weight.restype = ctypes.c_float
            lib.jcg_pair_weight.argtypes = [u64, i32, i32]
            lib.jcg_rmat_batch.restype = vp
            lib.jcg_rmat_batch.argtypes = [i32, vp, vp, vp, ctypes.c_int, u64, f64, f64, f64, vp, vp]
            lib.jcg_rmat_edges.restype = i64
            lib.jcg_rmat_edges.argtypes = [ctypes.c_int, i32, f64, f64, f64, f64, u64, vp, vp, i64]
            lib.jcio_load.restype = vp
            lib.jcio_load.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
            for name, rt in (("jcio_n", i64), ("jcio_m", i64), ("jcio_flags", u64)):
                getattr(lib, name).restype = rt
                getattr(lib, name).argtypes = [vp]
            lib.jcio_offsets.restype = ctypes.POINTER(i32)
            lib.jcio_offsets.argtypes = [vp]
            lib.jcio_indices.restype = ctypes.POINTER(i32)
            lib.jcio_indices.argtypes = [vp]
            lib.jcio_weights.restype = ctypes.POINTER(ctypes.c_float)
            lib.jcio_weights.argtypes = [vp]
            lib.jcio_free.restype = None
            lib.jcio_free.argtypes = [vp]
            lib.jcio_save.restype = ctypes.c_int
            lib.jcio_save.argtypes = [ctypes.c_char_p, i64, i64, vp, vp, vp, u64, ctypes.c_int]
            lib.jcg_stats.restype = None
            lib.jcg_stats.argtypes = [i32, vp, vp, ctypes.POINTER(i64)]
            _graph = lib
        return _graph


def _declare_jaccard(lib: ctypes.CDLL) -> None:
    vp, i32, i64, f32, u32, sz = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float,
                                  ctypes.c_uint, ctypes.c_size_t)
    I = ctypes.c_int
    sig = {
        "jc_jaccard": [vp, vp, vp, I, vp],
        "jc_jaccard_ex": [vp, vp, vp, I, u32, vp, sz, vp],
        "jc_jaccard_workspace_size": [i32, i32, I, I],
        "jc_jaccard_tuned": [vp, vp, vp, I, u32, vp, sz, vp, vp],
        "jc_jaccard_noalloc": [vp, vp, vp, I, u32, vp],
        "jc_jaccard_list": [vp, vp, i32, vp, vp, vp, vp],
        "jc_vertex_weights": [vp, vp, vp],
        "jc_vertex_weights_ex": [vp, vp, vp, sz, vp],
        "jc_vertex_weights_workspace_size": [i32],
        "jc_nonzero_count": [i32, vp, f32, ctypes.POINTER(i64), vp],
        "jc_fill": [i64, vp, f32, vp],
        "jc_rowsum": [i32, i32, vp, vp, vp, vp, vp],
        "jc_intersect": [i32, i32, vp, vp, vp, vp, vp, vp, vp],
        "jc_intersect_pairs": [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp],
        "jc_jw": [i32, vp, vp, vp, vp],
        "jc_ec_scan": [i32, i32, vp, vp, vp, vp],
        "jc_ec_unweighted": [i32, i32, vp, vp, vp, vp, vp],
        "jc_validate_csr": [vp, ctypes.POINTER(I), vp],
        "jc_csr_is_symmetric": [vp, ctypes.POINTER(I), vp],
        "jc_weighted_sum_mode": [vp, vp, ctypes.POINTER(I), vp],
        "jc_degree_stats": [vp, vp, vp],
        "jc_config_from_stats": [vp, i32, I, ctypes.POINTER(I), vp],
        "jc_select_config": [vp, I, ctypes.POINTER(I), vp, vp],
        "jc_csr_from_edges": [i32, i64, vp, vp, vp, I, vp, vp, vp, i64, ctypes.POINTER(i64), vp],
        "jc_profile_enable": [I],
        "jc_profile_reset": [],
        "jc_profile_query": [ctypes.c_char_p, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_double)],
        "jc_profile_names": [ctypes.c_char_p, sz],
        "jc_strerror": [I],
        "jc_last_hip_error": [],
        "jc_version": [],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = I
    lib.jc_jaccard_workspace_size.restype = sz
    lib.jc_vertex_weights_workspace_size.restype = sz
    lib.jc_strerror.restype = ctypes.c_char_p
    lib.jc_version.restype = ctypes.c_char_p


def load_jaccard_lib() -> ctypes.CDLL:
    """The HIP library; raises NativeLibraryMissing if it is not built."""
    global _jaccard
    with _lock:
        if _jaccard is None:
            # torch carries its own HIP runtime (libamdhip64.so, under a SONAME
            # other than ROCm's libamdhip64.so.7). Loaded first, it is the one
            # the library binds to; loaded after the library, the process would
            # hold two HIP runtimes, and torch's memory and streams would
            # belong to the other one.
            import torch  # noqa: F401
            lib = _load(JACCARD_LIB)
            _declare_jaccard(lib)
            _jaccard = lib
        return _jaccard
