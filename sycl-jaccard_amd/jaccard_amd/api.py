"""Python mirror of the reference's Jaccard interface over the C ABI.

Reference interface (paths relative to the reference root):
  * ``sygraph::jaccard<edge_centric,VT,ET,WT>(graph, [weights,] result, q)``
    standalone_algorithms.hpp:67-69, :86-87
  * ``sygraph::jaccard_list<VT,ET,WT>(graph, [weights,] num_pairs, first, second, result, q)``
    standalone_algorithms.hpp:109-112, :131-133
  * the kernel functors' ``invoke`` (jaccard.hpp:51-52, 104-107, 151-156,
    210-217, 235-237, 256-258, 280-283) and the driver's vertex-weight
    preamble (main.cpp:211-275)

Device memory and streams are torch tensors / torch streams (plumbing only);
all arithmetic runs in ``libjaccard_amd.so``. Errors raise ``JaccardError``
carrying the library's status code, where the reference printed a SYCL
exception and carried on.
"""
from __future__ import annotations

import ctypes
from typing import Optional

from ._native import load_jaccard_lib

EDGE_CENTRIC = 1
VERTEX_CENTRIC = 0
FLAG_ASSUME_SYMMETRIC = 1

ERR_INVALID_ARG = -1
ERR_INVALID_CSR = -2
ERR_OUT_OF_MEMORY = -3
ERR_HIP = -4
ERR_WORKSPACE = -5


class JaccardError(RuntimeError):
    def __init__(self, code: int, what: str):
        lib = load_jaccard_lib()
        msg = lib.jc_strerror(code).decode()
        if code == ERR_HIP:
            msg += f" (hipError {lib.jc_last_hip_error()})"
        super().__init__(f"{what}: {msg} [{code}]")
        self.code = code


class JcCsr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("m", ctypes.c_int32), ("offsets", ctypes.c_void_p),
                ("indices", ctypes.c_void_p), ("edge_w", ctypes.c_void_p)]


def _torch():
    import torch
    return torch


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _stream(stream) -> Optional[int]:
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def _check(code: int, what: str) -> None:
    if code != 0:
        raise JaccardError(code, what)


def _require(t, dtype, numel: int, name: str, device=None, optional: bool = False) -> None:
    """Raise ValueError unless ``t`` is a contiguous device tensor of ``dtype``
    with at least ``numel`` elements (on ``device`` when given). The library
    reads and writes through raw pointers, so a wrong dtype, a short buffer or
    host memory must be refused here, before any device access."""
    torch = _torch()
    if t is None:
        if optional:
            return
        raise ValueError(f"{name} is required")
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if t.numel() < numel:
        raise ValueError(f"{name} needs at least {numel} elements, got {t.numel()}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device (HIP) tensor, got one on {t.device}")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, the graph on {device}")


class DeviceCSR:
    """``GraphCSRView<int32_t,int32_t,float>`` resident on the GPU.

    ``offsets`` (int32[n+1]), ``indices`` (int32[m]) and optional ``edge_w``
    (float32[m]) are device tensors owned by this object.
    """

    def __init__(self, n: int, offsets, indices, edge_w=None):
        torch = _torch()
        for t, dt, name in ((offsets, torch.int32, "offsets"), (indices, torch.int32, "indices")):
            if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous int32 CUDA tensor")
        if offsets.numel() != n + 1:
            raise ValueError("offsets must have n+1 entries")
        if edge_w is not None and (edge_w.dtype != torch.float32 or edge_w.numel() != indices.numel()):
            raise ValueError("edge_w must be float32[m]")
        self.n = int(n)
        self.offsets = offsets
        self.indices = indices
        self.edge_w = edge_w

    @property
    def m(self) -> int:
        return int(self.indices.numel())

    @classmethod
    def from_host(cls, g, device="cuda") -> "DeviceCSR":
        torch = _torch()
        off = torch.from_numpy(g.offsets).to(device)
        ind = torch.from_numpy(g.indices).to(device) if g.m else torch.zeros(0, dtype=torch.int32, device=device)
        w = torch.from_numpy(g.weights).to(device) if g.weights is not None else None
        return cls(g.n, off, ind, w)

    def c_struct(self) -> JcCsr:
        return JcCsr(self.n, self.m, _ptr(self.offsets), _ptr(self.indices) if self.m else None,
                     _ptr(self.edge_w) if self.edge_w is not None and self.m else None)


def _empty_f32(n, device):
    torch = _torch()
    return torch.empty(max(n, 0), dtype=torch.float32, device=device)


class JcTuning(ctypes.Structure):
    """``jc_tuning``: launch parameters that do not change results."""
    _fields_ = [("small_max", ctypes.c_int32), ("small_blocks_per_cu", ctypes.c_int32),
                ("wide_index", ctypes.c_int32), ("bitmap_blocks_per_cu", ctypes.c_int32),
                ("hash_table", ctypes.c_int32), ("quot_max_disp", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 2)]


def jaccard(graph: DeviceCSR, result=None, weights=None, edge_centric: bool = True, stream=None,
            flags: int = 0, workspace=None, tuning=None):
    """Jaccard coefficient of every stored edge (``sygraph::jaccard``).

    ``weights``: float32[n] vertex weights, or None for the unweighted form.
    ``result``: float32[m] output (allocated when None). Returns ``result``.
    ``tuning``: dict of ``jc_tuning`` fields (``jc_jaccard_tuned``), or None.
    """
    torch = _torch()
    lib = load_jaccard_lib()
    dev = graph.offsets.device
    if result is None:
        result = _empty_f32(graph.m, dev)
    _require(weights, torch.float32, graph.n, "weights", dev, optional=True)
    _require(result, torch.float32, graph.m, "result", dev)
    if workspace is not None:
        _require(workspace, workspace.dtype, 0, "workspace", dev)
    g = graph.c_struct()
    ws_ptr, ws_bytes = (None, 0) if workspace is None else (workspace.data_ptr(), workspace.numel() * workspace.element_size())
    algo = EDGE_CENTRIC if edge_centric else VERTEX_CENTRIC
    out = _ptr(result) if graph.m else None
    if tuning is None:
        _check(lib.jc_jaccard_ex(ctypes.byref(g), _ptr(weights), out, algo, flags, ws_ptr, ws_bytes,
                                 _stream(stream)), "jc_jaccard")
    else:
        t = JcTuning(**tuning)
        _check(lib.jc_jaccard_tuned(ctypes.byref(g), _ptr(weights), out, algo, flags, ws_ptr, ws_bytes,
                                    ctypes.byref(t), _stream(stream)), "jc_jaccard_tuned")
    return result


ERR_SCRATCH_LIMIT = -6


def jaccard_noalloc(graph: DeviceCSR, result=None, weights=None, edge_centric: bool = True, stream=None,
                    flags: int = 0):
    """``jaccard`` through ``jc_jaccard_noalloc``: the library allocates no
    device memory (its workspace is a fixed per-device array of
    ``NOALLOC_SCRATCH_BYTES``); graphs that need more raise JaccardError
    ERR_SCRATCH_LIMIT. Overlapping calls on one device must be serialized."""
    torch = _torch()
    lib = load_jaccard_lib()
    dev = graph.offsets.device
    if result is None:
        result = _empty_f32(graph.m, dev)
    _require(weights, torch.float32, graph.n, "weights", dev, optional=True)
    _require(result, torch.float32, graph.m, "result", dev)
    g = graph.c_struct()
    algo = EDGE_CENTRIC if edge_centric else VERTEX_CENTRIC
    _check(lib.jc_jaccard_noalloc(ctypes.byref(g), _ptr(weights), _ptr(result) if graph.m else None, algo, flags,
                                  _stream(stream)), "jc_jaccard_noalloc")
    return result


NOALLOC_SCRATCH_BYTES = 32 << 20


def workspace_size(n: int, m: int, weighted: bool, edge_centric: bool) -> int:
    return int(load_jaccard_lib().jc_jaccard_workspace_size(n, m, int(weighted), int(edge_centric)))


def jaccard_list(graph: DeviceCSR, first, second, result=None, weights=None, stream=None):
    """Jaccard coefficient of arbitrary vertex pairs (``sygraph::jaccard_list``)."""
    torch = _torch()
    lib = load_jaccard_lib()
    dev = graph.offsets.device
    if isinstance(first, torch.Tensor) and isinstance(second, torch.Tensor) and first.numel() != second.numel():
        raise ValueError("first and second must have the same length")
    _require(first, torch.int32, 0, "first", dev)
    k = int(first.numel())
    _require(second, torch.int32, k, "second", dev)
    if result is None:
        result = _empty_f32(k, dev)
    _require(result, torch.float32, k, "result", dev)
    _require(weights, torch.float32, graph.n, "weights", dev, optional=True)
    g = graph.c_struct()
    _check(lib.jc_jaccard_list(ctypes.byref(g), _ptr(weights), k, _ptr(first), _ptr(second), _ptr(result),
                               _stream(stream)), "jc_jaccard_list")
    return result


def vertex_weights(graph: DeviceCSR, out=None, stream=None):
    """W[k] = sum of row k's edge weights (the reference driver's preamble)."""
    lib = load_jaccard_lib()
    if out is None:
        out = _empty_f32(graph.n, graph.offsets.device)
    _require(out, _torch().float32, graph.n, "out", graph.offsets.device)
    g = graph.c_struct()
    _check(lib.jc_vertex_weights(ctypes.byref(g), _ptr(out), _stream(stream)), "jc_vertex_weights")
    return out


def vertex_weights_ex(graph: DeviceCSR, out, workspace, stream=None):
    """vertex_weights with a caller workspace (allocates nothing)."""
    lib = load_jaccard_lib()
    _require(out, _torch().float32, graph.n, "out", graph.offsets.device)
    _require(workspace, workspace.dtype, 0, "workspace", graph.offsets.device)
    g = graph.c_struct()
    _check(lib.jc_vertex_weights_ex(ctypes.byref(g), _ptr(out), workspace.data_ptr(),
                                    workspace.numel() * workspace.element_size(), _stream(stream)),
           "jc_vertex_weights_ex")
    return out


def vertex_weights_workspace_size(n: int) -> int:
    return int(load_jaccard_lib().jc_vertex_weights_workspace_size(n))


def nonzero_count(result, thresh: float = 1e-5, stream=None) -> int:
    """Entries above ``thresh`` (the EC path's "non zero pairs" line)."""
    lib = load_jaccard_lib()
    _require(result, _torch().float32, 0, "result")
    c = ctypes.c_int64(0)
    _check(lib.jc_nonzero_count(int(result.numel()), _ptr(result), float(thresh), ctypes.byref(c),
                                _stream(stream)), "jc_nonzero_count")
    return int(c.value)


# ---- kernel level ----------------------------------------------------------

def fill(x, value: float, stream=None):
    _require(x, _torch().float32, 0, "x")
    _check(load_jaccard_lib().jc_fill(int(x.numel()), _ptr(x), float(value), _stream(stream)), "jc_fill")
    return x


def rowsum(graph: DeviceCSR, work, v=None, stream=None):
    f32, dev = _torch().float32, graph.offsets.device
    _require(work, f32, graph.n, "work", dev)
    _require(v, f32, graph.n, "v", dev, optional=True)
    _check(load_jaccard_lib().jc_rowsum(graph.n, graph.m, _ptr(graph.offsets), _ptr(graph.indices), _ptr(v),
                                        _ptr(work), _stream(stream)), "jc_rowsum")
    return work


def intersect(graph: DeviceCSR, work, weight_i, weight_s, v=None, stream=None):
    f32, dev = _torch().float32, graph.offsets.device
    _require(work, f32, graph.n, "work", dev)
    _require(weight_i, f32, graph.m, "weight_i", dev)
    _require(weight_s, f32, graph.m, "weight_s", dev)
    _require(v, f32, graph.n, "v", dev, optional=True)
    _check(load_jaccard_lib().jc_intersect(graph.n, graph.m, _ptr(graph.offsets), _ptr(graph.indices), _ptr(v),
                                           _ptr(work), _ptr(weight_i), _ptr(weight_s), _stream(stream)),
           "jc_intersect")


def intersect_pairs(graph: DeviceCSR, first, second, work, weight_i, weight_s, v=None, stream=None):
    torch, dev = _torch(), graph.offsets.device
    _require(first, torch.int32, 0, "first", dev)
    k = int(first.numel())
    _require(second, torch.int32, k, "second", dev)
    _require(work, torch.float32, graph.n, "work", dev)
    _require(weight_i, torch.float32, k, "weight_i", dev)
    _require(weight_s, torch.float32, k, "weight_s", dev)
    _require(v, torch.float32, graph.n, "v", dev, optional=True)
    _check(load_jaccard_lib().jc_intersect_pairs(graph.n, int(first.numel()), _ptr(graph.offsets),
                                                 _ptr(graph.indices), _ptr(first), _ptr(second), _ptr(v),
                                                 _ptr(work), _ptr(weight_i), _ptr(weight_s), _stream(stream)),
           "jc_intersect_pairs")


def jw(weight_i, weight_s, weight_j, stream=None):
    f32 = _torch().float32
    _require(weight_i, f32, 0, "weight_i")
    _require(weight_s, f32, weight_i.numel(), "weight_s", weight_i.device)
    _require(weight_j, f32, weight_i.numel(), "weight_j", weight_i.device)
    _check(load_jaccard_lib().jc_jw(int(weight_i.numel()), _ptr(weight_i), _ptr(weight_s), _ptr(weight_j),
                                    _stream(stream)), "jc_jw")
    return weight_j


def ec_scan(graph: DeviceCSR, dest_ind, weight_j, stream=None):
    torch, dev = _torch(), graph.offsets.device
    _require(dest_ind, torch.int32, graph.m, "dest_ind", dev)
    _require(weight_j, torch.float32, graph.m, "weight_j", dev)
    _check(load_jaccard_lib().jc_ec_scan(graph.m, graph.n, _ptr(graph.offsets), _ptr(dest_ind), _ptr(weight_j),
                                         _stream(stream)), "jc_ec_scan")


def ec_unweighted(graph: DeviceCSR, dest_ind, weight_j, stream=None):
    torch, dev = _torch(), graph.offsets.device
    _require(dest_ind, torch.int32, graph.m, "dest_ind", dev, optional=True)
    _require(weight_j, torch.float32, graph.m, "weight_j", dev)
    _check(load_jaccard_lib().jc_ec_unweighted(graph.m, graph.n, _ptr(graph.offsets), _ptr(graph.indices),
                                               _ptr(dest_ind), _ptr(weight_j), _stream(stream)),
           "jc_ec_unweighted")
    return weight_j


def validate_csr(graph: DeviceCSR, stream=None) -> int:
    """0 when valid, else a bit mask: 1 offsets, 2 column range, 4 row order."""
    lib = load_jaccard_lib()
    what = ctypes.c_int(0)
    g = graph.c_struct()
    code = lib.jc_validate_csr(ctypes.byref(g), ctypes.byref(what), _stream(stream))
    if code not in (0, ERR_INVALID_CSR):
        raise JaccardError(code, "jc_validate_csr")
    return int(what.value)


def csr_from_edges(n: int, src, dst, edge_weights=None, symmetrize: bool = True, weighted_out: bool = None,
                   stream=None) -> DeviceCSR:
    """Device graph ingest (jc_csr_from_edges): src/dst/edge_weights are device
    int32/float32 tensors of one directed edge each; returns a sorted,
    duplicate-free CSR without self loops (reverse edges added when
    symmetrize). Weights are produced when edge_weights is given (or
    weighted_out)."""
    torch = _torch()
    lib = load_jaccard_lib()
    _require(src, torch.int32, 0, "src")
    e = int(src.numel())
    dev = src.device
    _require(dst, torch.int32, e, "dst", dev)
    _require(edge_weights, torch.float32, e, "edge_weights", dev, optional=True)
    cap = max(e * (2 if symmetrize else 1), 1)
    off = torch.empty(n + 1, dtype=torch.int32, device=dev)
    ind = torch.empty(cap, dtype=torch.int32, device=dev)
    want_w = edge_weights is not None if weighted_out is None else weighted_out
    w = torch.empty(cap, dtype=torch.float32, device=dev) if want_w else None
    m = ctypes.c_int64(0)
    _check(lib.jc_csr_from_edges(n, e, src.data_ptr(), dst.data_ptr(),
                                 None if edge_weights is None else edge_weights.data_ptr(), 1 if symmetrize else 0,
                                 off.data_ptr(), ind.data_ptr(), None if w is None else w.data_ptr(), cap,
                                 ctypes.byref(m), _stream(stream)), "jc_csr_from_edges")
    mm = int(m.value)
    return DeviceCSR(n, off, ind[:mm], None if w is None else w[:mm])


SUM_FP64 = 0
SUM_DOUBLE_DOUBLE = 1


def weighted_sum_mode(graph: DeviceCSR, weights, stream=None) -> int:
    """How weighted calls add ``weights`` on ``graph`` (jc_weighted_sum_mode):
    SUM_FP64 when plain fp64 sums are exact, else SUM_DOUBLE_DOUBLE. Synchronous."""
    _require(weights, _torch().float32, graph.n, "weights", graph.offsets.device)
    mode = ctypes.c_int(-1)
    g = graph.c_struct()
    _check(load_jaccard_lib().jc_weighted_sum_mode(ctypes.byref(g), _ptr(weights), ctypes.byref(mode),
                                                   _stream(stream)), "jc_weighted_sum_mode")
    return int(mode.value)


class JcGraphStats(ctypes.Structure):
    """``jc_graph_stats``: degree statistics of a graph (jc_degree_stats)."""
    _fields_ = [("rows", ctypes.c_int64 * 32), ("entries", ctypes.c_int64 * 32), ("isolated", ctypes.c_int64),
                ("long_rows", ctypes.c_int64), ("long_bitmap_rows", ctypes.c_int64 * 2),
                ("max_degree", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("bitmap_table_bytes", ctypes.c_int64 * 2)]

    def as_dict(self) -> dict:
        return {"rows": list(self.rows), "entries": list(self.entries), "isolated": int(self.isolated),
                "long_rows": int(self.long_rows), "long_bitmap_rows": list(self.long_bitmap_rows),
                "max_degree": int(self.max_degree), "bitmap_table_bytes": list(self.bitmap_table_bytes)}


_TUNING_FIELDS = ("small_max", "small_blocks_per_cu", "wide_index", "bitmap_blocks_per_cu", "hash_table",
                  "quot_max_disp")


def degree_stats(graph: DeviceCSR, stream=None) -> JcGraphStats:
    """Degree histogram and long-row facts of ``graph`` (jc_degree_stats). Synchronous."""
    st = JcGraphStats()
    g = graph.c_struct()
    _check(load_jaccard_lib().jc_degree_stats(ctypes.byref(g), ctypes.byref(st), _stream(stream)),
           "jc_degree_stats")
    return st


def config_from_stats(stats: JcGraphStats, n: int, weighted: bool):
    """(edge_centric, tuning dict) that jc_config_from_stats picks (host only)."""
    ec, t = ctypes.c_int(-1), JcTuning()
    _check(load_jaccard_lib().jc_config_from_stats(ctypes.byref(stats), int(n), int(bool(weighted)),
                                                   ctypes.byref(ec), ctypes.byref(t)), "jc_config_from_stats")
    return ec.value == EDGE_CENTRIC, {f: int(getattr(t, f)) for f in _TUNING_FIELDS}


def select_config(graph: DeviceCSR, weighted: bool, stream=None):
    """(edge_centric, tuning dict) chosen from ``graph``'s degree statistics
    (jc_select_config), for ``jaccard(..., edge_centric=..., tuning=...)``. Synchronous."""
    ec, t = ctypes.c_int(-1), JcTuning()
    g = graph.c_struct()
    _check(load_jaccard_lib().jc_select_config(ctypes.byref(g), int(bool(weighted)), ctypes.byref(ec),
                                               ctypes.byref(t), _stream(stream)), "jc_select_config")
    return ec.value == EDGE_CENTRIC, {f: int(getattr(t, f)) for f in _TUNING_FIELDS}


def is_symmetric(graph: DeviceCSR, stream=None) -> bool:
    lib = load_jaccard_lib()
    sym = ctypes.c_int(0)
    g = graph.c_struct()
    _check(lib.jc_csr_is_symmetric(ctypes.byref(g), ctypes.byref(sym), _stream(stream)), "jc_csr_is_symmetric")
    return bool(sym.value)


def profile_enable(on: bool = True) -> None:
    """Per-kernel event timing (the reference's EVENT_PROFILE counterpart)."""
    load_jaccard_lib().jc_profile_enable(1 if on else 0)


def profile_reset() -> None:
    load_jaccard_lib().jc_profile_reset()


def profile_query(kernel: Optional[str] = None):
    """(launch count, total ms) of one kernel name, or of all when None. Synchronous."""
    c, t = ctypes.c_int(0), ctypes.c_double(0.0)
    _check(load_jaccard_lib().jc_profile_query(kernel.encode() if kernel else None, ctypes.byref(c),
                                               ctypes.byref(t)), "jc_profile_query")
    return int(c.value), float(t.value)


def profile_names():
    buf = ctypes.create_string_buffer(4096)
    load_jaccard_lib().jc_profile_names(buf, 4096)
    return [x for x in buf.value.decode().split(",") if x]


def version() -> str:
    return load_jaccard_lib().jc_version().decode()
