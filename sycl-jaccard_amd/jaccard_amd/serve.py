"""Serving: timed Jaccard requests against graphs resident in HBM.

``Server`` wraps the native serving runtime (jc_server_* in jaccard_amd.h):
the graphs are uploaded once, packed on the device into one arena (their
disjoint union), and a stream of requests -- whole-graph edge-centric Jaccard
or the Jaccard of k vertex pairs -- is replayed with its arrival times. In
continuous mode requests that arrive while others run are admitted into the
next batch as soon as a slot (HIP stream) and HBM-budget memory are free; in
FIFO mode each request is served alone with the plain entry points
(jc_jaccard / jc_jaccard_list), the per-request baseline.

No reference counterpart: the reference computes one graph per call
(jaccard.cpp:926-955).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

from . import api, graphs
from ._native import load_jaccard_lib

CONTINUOUS = 0
FIFO = 1
REQ_FULL = 0
REQ_PAIRS = 1


class JcServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int32), ("max_batch_requests", ctypes.c_int32),
                ("max_batch_edges", ctypes.c_int64), ("max_batch_pairs", ctypes.c_int64),
                ("hbm_budget_bytes", ctypes.c_size_t), ("mode", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 5)]


class JcRequest(ctypes.Structure):
    _fields_ = [("graph", ctypes.c_int32), ("kind", ctypes.c_int32), ("num_pairs", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("first", ctypes.c_void_p), ("second", ctypes.c_void_p),
                ("result", ctypes.c_void_p), ("arrival", ctypes.c_double)]


class JcRequestStats(ctypes.Structure):
    _fields_ = [("admitted", ctypes.c_double), ("completed", ctypes.c_double),
                ("batch", ctypes.c_int32), ("status", ctypes.c_int32)]


class JcServerRunInfo(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_int32), ("max_in_flight", ctypes.c_int32),
                ("peak_budget_bytes", ctypes.c_size_t), ("wall_seconds", ctypes.c_double)]


def _bind(lib):
    if getattr(lib, "_jc_serve_bound", False):
        return lib
    vp = ctypes.c_void_p
    lib.jc_server_create.argtypes = [ctypes.c_int32, vp, vp, ctypes.POINTER(JcServerConfig), ctypes.POINTER(vp)]
    lib.jc_server_create.restype = ctypes.c_int
    lib.jc_server_destroy.argtypes = [vp]
    lib.jc_server_destroy.restype = ctypes.c_int
    lib.jc_server_arena_bytes.argtypes = [vp]
    lib.jc_server_arena_bytes.restype = ctypes.c_size_t
    lib.jc_server_run.argtypes = [vp, ctypes.POINTER(JcRequest), ctypes.c_int32, ctypes.POINTER(JcRequestStats),
                                  ctypes.POINTER(JcServerRunInfo)]
    lib.jc_server_run.restype = ctypes.c_int
    lib._jc_serve_bound = True
    return lib


@dataclasses.dataclass
class Request:
    graph: int
    kind: int                        # REQ_FULL or REQ_PAIRS
    arrival: float                   # seconds after the start of the run
    first: Optional[np.ndarray] = None   # int32 local ids (pairs)
    second: Optional[np.ndarray] = None

    @property
    def num_pairs(self) -> int:
        return 0 if self.first is None else int(self.first.size)


class Server:
    """Resident graphs + native serving runtime (one mode per server)."""

    def __init__(self, gs: Sequence[graphs.CSRGraph], device="cuda", weighted: bool = True, mode: int = CONTINUOUS,
                 slots: int = 4, max_batch_requests: int = 256, max_batch_edges: int = 1 << 26,
                 max_batch_pairs: int = 1 << 20, hbm_budget_bytes: int = 256 << 20, device_graphs=None):
        torch = api._torch()
        self.lib = _bind(load_jaccard_lib())
        self.device = torch.device(device)
        self.weighted = weighted
        self.mode = mode
        self.host = list(gs)
        # device CSRs (local ids) and vertex weights, shareable between servers
        if device_graphs is None:
            device_graphs = upload(gs, self.device, weighted)
        self.dev = device_graphs
        n = len(gs)
        csrs = (api.JcCsr * n)(*[d[0].c_struct() for d in self.dev])
        self._csrs = csrs
        self._wptrs = None
        wp = None
        if weighted:
            self._wptrs = (ctypes.c_void_p * n)(*[d[1].data_ptr() for d in self.dev])
            wp = ctypes.cast(self._wptrs, ctypes.c_void_p)
        cfg = JcServerConfig(slots, max_batch_requests, max_batch_edges, max_batch_pairs, hbm_budget_bytes, mode)
        h = ctypes.c_void_p()
        torch.cuda.synchronize(self.device)
        code = self.lib.jc_server_create(n, ctypes.cast(csrs, ctypes.c_void_p), wp, ctypes.byref(cfg), ctypes.byref(h))
        if code != 0:
            raise api.JaccardError(code, "jc_server_create")
        self.h = h

    def arena_bytes(self) -> int:
        return int(self.lib.jc_server_arena_bytes(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.jc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, reqs: List[Request], results: Sequence):
        """Replays reqs; results[i] is a device float32 tensor of m (full) or
        k (pairs) entries. Returns (stats structured array, info dict)."""
        torch = api._torch()
        cnt = len(reqs)
        arr = (JcRequest * cnt)()
        keep = []
        for i, q in enumerate(reqs):
            r = arr[i]
            r.graph, r.kind, r.arrival = q.graph, q.kind, float(q.arrival)
            r.result = results[i].data_ptr() if results[i] is not None else None
            if q.kind == REQ_PAIRS:
                f = np.ascontiguousarray(q.first, np.int32)
                s = np.ascontiguousarray(q.second, np.int32)
                keep += [f, s]
                r.num_pairs = f.size
                r.first, r.second = f.ctypes.data, s.ctypes.data
        st = (JcRequestStats * cnt)()
        info = JcServerRunInfo()
        torch.cuda.synchronize(self.device)
        code = self.lib.jc_server_run(self.h, arr, cnt, st, ctypes.byref(info))
        if code != 0:
            raise api.JaccardError(code, "jc_server_run")
        out = np.zeros(cnt, dtype=[("admitted", "f8"), ("completed", "f8"), ("batch", "i4"), ("status", "i4")])
        for i in range(cnt):
            out[i] = (st[i].admitted, st[i].completed, st[i].batch, st[i].status)
        return out, {"batches": info.batches, "max_in_flight": info.max_in_flight,
                     "peak_budget_bytes": int(info.peak_budget_bytes), "wall_seconds": info.wall_seconds}


def upload(gs: Sequence[graphs.CSRGraph], device, weighted: bool):
    """Device CSR (local ids) and vertex weights W (row sums of edge weights) per graph."""
    torch = api._torch()
    outs = []
    for g in gs:
        dg = api.DeviceCSR.from_host(g, device)
        W = None
        if weighted:
            W = torch.empty(max(g.n, 1), dtype=torch.float32, device=device)
            api.vertex_weights(dg, out=W)
        outs.append((dg, W))
    torch.cuda.synchronize(device)
    return outs


def serving_graphs(count: int = 64, scale_min: int = 14, scale_max: int = 20, edge_factor: int = 16,
                   seed: int = 7, weighted: bool = True) -> List[graphs.CSRGraph]:
    """count RMAT graphs (a/b/c = .57/.19/.19), scales cycling scale_min..scale_max,
    float32 edge weights in (0, 1], generated in parallel."""
    from . import batch
    scales = [scale_min + (i % (scale_max - scale_min + 1)) for i in range(count)]
    ns = [1 << s for s in scales]
    seeds = [seed * 1000003 + i for i in range(count)]
    b = batch.rmat_batch(ns, [float(edge_factor)] * count, seeds, weighted)
    out = []
    for i in range(count):
        g = b.member(i)
        g.name = f"rmat{scales[i]}_{i}"
        out.append(g)
    return out


def request_stream(gs: Sequence[graphs.CSRGraph], count: int, qps: float, p_full: float = 0.5,
                   kmax: int = 65536, seed: int = 11) -> List[Request]:
    """Poisson arrivals at rate qps (qps <= 0: all at t = 0). Each request picks a
    graph uniformly; it is a whole-graph query with probability p_full, else k
    vertex pairs with k log-uniform in [1, kmax], half of them stored edges and
    half uniform random vertex pairs."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for _ in range(count):
        if qps > 0:
            t += rng.exponential(1.0 / qps)
        gi = int(rng.integers(0, len(gs)))
        g = gs[gi]
        if rng.random() < p_full:
            reqs.append(Request(gi, REQ_FULL, t))
            continue
        k = int(min(kmax, math.floor(math.exp(rng.uniform(0.0, math.log(kmax + 1))))))
        k = max(k, 1)
        ne = k // 2 if g.m > 0 else 0
        e = rng.integers(0, max(g.m, 1), ne)
        rows = np.searchsorted(g.offsets, e, side="right") - 1
        first = np.concatenate([rows, rng.integers(0, g.n, k - ne)]).astype(np.int32)
        second = np.concatenate([g.indices[e], rng.integers(0, g.n, k - ne)]).astype(np.int32)
        perm = rng.permutation(k)
        reqs.append(Request(gi, REQ_PAIRS, t, first[perm], second[perm]))
    return reqs


def result_buffers(gs: Sequence[graphs.CSRGraph], reqs: List[Request], device, share_full: bool = False):
    """Device result tensors; share_full gives requests on the same graph one buffer."""
    torch = api._torch()
    per_graph = {}
    out = []
    for q in reqs:
        if q.kind == REQ_FULL:
            if share_full:
                if q.graph not in per_graph:
                    per_graph[q.graph] = torch.empty(max(gs[q.graph].m, 1), dtype=torch.float32, device=device)
                out.append(per_graph[q.graph])
            else:
                out.append(torch.empty(max(gs[q.graph].m, 1), dtype=torch.float32, device=device))
        else:
            out.append(torch.empty(max(q.num_pairs, 1), dtype=torch.float32, device=device))
    return out


def latency_summary(reqs: List[Request], st) -> dict:
    arr = np.array([q.arrival for q in reqs])
    lat = st["completed"] - arr
    kinds = np.array([q.kind for q in reqs])
    wall = float(st["completed"].max() - arr.min()) if len(reqs) else 0.0
    d = {"requests": len(reqs), "wall_s": wall, "requests_per_s": len(reqs) / wall if wall > 0 else None,
         "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
         "mean_ms": float(lat.mean() * 1e3)}
    for k, name in ((REQ_FULL, "full"), (REQ_PAIRS, "pairs")):
        sel = lat[kinds == k]
        if sel.size:
            d[f"{name}_p50_ms"] = float(np.percentile(sel, 50) * 1e3)
            d[f"{name}_p99_ms"] = float(np.percentile(sel, 99) * 1e3)
    return d
