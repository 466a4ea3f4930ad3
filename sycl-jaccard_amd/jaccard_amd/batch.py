"""Batched Jaccard over thousands of graphs in one launch.

A batch of independent CSR graphs, packed with global vertex ids, is one
block-diagonal CSR (their disjoint union): no row references another graph's
vertices, so one owner-engine pass over the union computes every graph, and
each graph's outputs are bit-identical to a separate call (unweighted counts
are exact integers; weighted sums are exact double-double, so the reduction
order -- which the union changes -- does not matter).

``BatchRunner`` keeps the union resident in HBM with a preallocated
workspace, picks launch parameters with a small autotuner (degree-bin
threshold between the one-wave-per-row kernel and the workgroup kernel, grid
of the former, one or two workgroups per CU for the bitmap-table kernel; none
of them changes results), and captures the whole step in
a HIP graph that is replayed per batch. ``PlainRunner`` is the comparison:
one ``jc_jaccard`` call per graph, back to back on one stream.

No reference counterpart: the reference computes one graph per call
(jaccard.cpp:926-955).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import time
from typing import List, Optional, Sequence

import numpy as np

from . import api, graphs
from ._native import load_graph_lib, load_jaccard_lib


@dataclasses.dataclass
class GraphBatch:
    union: graphs.CSRGraph  # block-diagonal CSR with global vertex ids
    vbase: np.ndarray       # int64[count+1]: first vertex of graph i
    ebase: np.ndarray       # int64[count+1]: first stored edge of graph i
    ids: np.ndarray         # batch-wide graph ids of the members

    @property
    def count(self) -> int:
        return int(self.vbase.size - 1)

    def member(self, i: int) -> graphs.CSRGraph:
        """Graph i with local vertex ids (host copy)."""
        vb, ve = int(self.vbase[i]), int(self.vbase[i + 1])
        eb, ee = int(self.ebase[i]), int(self.ebase[i + 1])
        off = (self.union.offsets[vb:ve + 1] - eb).astype(np.int32)
        ind = (self.union.indices[eb:ee] - vb).astype(np.int32)
        w = None if self.union.weights is None else self.union.weights[eb:ee].copy()
        return graphs.CSRGraph(n=ve - vb, offsets=off, indices=ind, weights=w, name=f"member{i}")


def rmat_batch(ns: Sequence[int], efs: Sequence[float], seeds: Sequence[int], weighted: bool,
               ids: Optional[Sequence[int]] = None) -> GraphBatch:
    """Generates RMAT graphs (a/b/c = .57/.19/.19) straight into a union CSR."""
    lib = load_graph_lib()
    cnt = len(ns)
    n = np.ascontiguousarray(ns, np.int32)
    ef = np.ascontiguousarray(efs, np.float64)
    sd = np.ascontiguousarray(seeds, np.uint64)
    vbase = np.zeros(cnt + 1, np.int64)
    ebase = np.zeros(cnt + 1, np.int64)
    h = lib.jcg_rmat_batch(cnt, n.ctypes.data, ef.ctypes.data, sd.ctypes.data, 1 if weighted else 0,
                           ctypes.c_uint64(graphs.DEFAULT_WSEED_XOR), graphs.RMAT_A, graphs.RMAT_B, graphs.RMAT_C,
                           vbase.ctypes.data, ebase.ctypes.data)
    u = graphs._take(h, f"rmat_batch{cnt}")
    return GraphBatch(union=u, vbase=vbase, ebase=ebase,
                      ids=np.arange(cnt) if ids is None else np.asarray(ids))


def batch_specs(count: int = 4096, seed: int = 1, nmin: int = 1 << 8, nmax: int = 1 << 16,
                efmin: int = 4, efmax: int = 32):
    """The batched workload: vertex counts log-uniform in [nmin, nmax], integer
    edge factors uniform in [efmin, efmax], graph i seeded by (seed, i), even i
    weighted. Deterministic."""
    rng = np.random.default_rng(seed)
    ns = np.floor(np.exp(rng.uniform(math.log(nmin), math.log(nmax), count))).astype(np.int64)
    ns = np.clip(ns, nmin, nmax).astype(np.int32)
    efs = rng.integers(efmin, efmax + 1, count).astype(np.float64)
    # seed * golden-ratio constant mod 2^64 (Python integers: no overflow warning)
    base = np.uint64((int(seed) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)
    seeds = np.arange(count, dtype=np.uint64) + base
    weighted = (np.arange(count) % 2) == 0
    return ns, efs, seeds, weighted


def make_batches(count: int = 4096, seed: int = 1, **kw):
    """(unweighted batch, weighted batch) of the batched workload."""
    ns, efs, seeds, weighted = batch_specs(count, seed, **kw)
    idx_u = np.nonzero(~weighted)[0]
    idx_w = np.nonzero(weighted)[0]
    bu = rmat_batch(ns[idx_u], efs[idx_u], seeds[idx_u], False, ids=idx_u)
    bw = rmat_batch(ns[idx_w], efs[idx_w], seeds[idx_w], True, ids=idx_w)
    return bu, bw


JcTuning = api.JcTuning


def _tuned(dg, weights, out, ws, tuning, stream):
    lib = load_jaccard_lib()
    g = dg.c_struct()
    t = None if tuning is None else ctypes.byref(
        JcTuning(small_max=tuning[0], small_blocks_per_cu=tuning[1], bitmap_blocks_per_cu=tuning[2]))
    code = lib.jc_jaccard_tuned(ctypes.byref(g), None if weights is None else weights.data_ptr(), out.data_ptr(),
                                1, api.FLAG_ASSUME_SYMMETRIC, ws.data_ptr(), ws.numel(), t, api._stream(stream))
    if code != 0:
        raise api.JaccardError(code, "jc_jaccard_tuned")


class BatchRunner:
    """Device-resident union batches, one captured HIP graph per step."""

    # (small_max, small_blocks_per_cu, bitmap_blocks_per_cu)
    TUNING_SPACE = [(128, 16, 1), (64, 16, 1), (128, 32, 1), (64, 32, 1), (128, 32, 2)]

    def __init__(self, batches: List[GraphBatch], device="cuda"):
        torch = api._torch()
        self.device = torch.device(device)
        self.items = []
        for b in batches:
            dg = api.DeviceCSR.from_host(b.union, self.device)
            weighted = b.union.weights is not None
            out = torch.empty(dg.m, dtype=torch.float32, device=self.device)
            W = torch.empty(dg.n, dtype=torch.float32, device=self.device) if weighted else None
            ws = torch.empty(api.workspace_size(dg.n, dg.m, weighted, True), dtype=torch.uint8, device=self.device)
            wsv = (torch.empty(api.vertex_weights_workspace_size(dg.n), dtype=torch.uint8, device=self.device)
                   if weighted else None)
            self.items.append(dict(batch=b, dg=dg, out=out, W=W, ws=ws, wsv=wsv))
        self.tuning = [None] * len(self.items)
        self.graph = None
        self.stream = torch.cuda.Stream(device=self.device)

    def _run_item(self, it, tuning, stream):
        if it["W"] is not None:
            api.vertex_weights_ex(it["dg"], it["W"], it["wsv"], stream=stream)
        _tuned(it["dg"], it["W"], it["out"], it["ws"], tuning, stream)

    def step(self, stream=None):
        for it, t in zip(self.items, self.tuning):
            self._run_item(it, t, stream)

    def select(self):
        """Per batch, the launch parameters jc_select_config picks from the
        union's degree statistics (no timing runs)."""
        for i, it in enumerate(self.items):
            _, t = api.select_config(it["dg"], it["W"] is not None)
            self.tuning[i] = (t["small_max"], t["small_blocks_per_cu"], t["bitmap_blocks_per_cu"])
        return list(self.tuning)

    def autotune(self, reps: int = 2):
        """Per batch, the fastest launch parameters of TUNING_SPACE and of the
        degree-statistics choice (select)."""
        torch = api._torch()
        report = []
        chosen = self.select()
        for i, it in enumerate(self.items):
            best = None
            space = list(self.TUNING_SPACE)
            if chosen[i] not in space:
                space.insert(0, chosen[i])
            for cand in space:
                self._run_item(it, cand, None)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    self._run_item(it, cand, None)
                torch.cuda.synchronize()
                dt = (time.perf_counter() - t0) / reps
                report.append((i, cand, dt))
                if best is None or dt < best[1]:
                    best = (cand, dt)
            self.tuning[i] = best[0]
        return report

    def capture(self):
        """Records one step into a HIP graph (stream capture)."""
        torch = api._torch()
        self.step(None)  # warm: module load, attributes
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self.step(torch.cuda.current_stream())
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()

    def member_output(self, k: int, i: int):
        """Output slice of member i of batch k (device tensor view)."""
        it = self.items[k]
        b = it["batch"]
        return it["out"][int(b.ebase[i]):int(b.ebase[i + 1])]


class PlainRunner:
    """One drop-in jc_jaccard call per graph (local ids), back to back: for a
    weighted graph jc_vertex_weights then jc_jaccard, as the reference driver
    does (main.cpp:211-275). The calls go straight to the C ABI with their
    arguments prepared beforehand, so the loop costs the library's own
    per-call work (launches, stream-ordered temporaries) plus one ctypes call
    each, not Python argument checks."""

    def __init__(self, batches: List[GraphBatch], device="cuda"):
        torch = api._torch()
        self.device = torch.device(device)
        self.graphs = []   # (batch index, member index, DeviceCSR, out, W)
        for k, b in enumerate(batches):
            uo = torch.from_numpy(b.union.offsets).to(self.device)
            ui = torch.from_numpy(b.union.indices).to(self.device)
            uw = None if b.union.weights is None else torch.from_numpy(b.union.weights).to(self.device)
            for i in range(b.count):
                vb, ve = int(b.vbase[i]), int(b.vbase[i + 1])
                eb, ee = int(b.ebase[i]), int(b.ebase[i + 1])
                off = (uo[vb:ve + 1] - eb).contiguous()
                ind = (ui[eb:ee] - vb).contiguous()
                w = None if uw is None else uw[eb:ee].contiguous()
                dg = api.DeviceCSR(ve - vb, off, ind, w)
                out = torch.empty(ee - eb, dtype=torch.float32, device=self.device)
                W = torch.empty(ve - vb, dtype=torch.float32, device=self.device) if w is not None else None
                self.graphs.append((k, i, dg, out, W))
            del uo, ui, uw
        # interleave as generated (global graph id order)
        self.graphs.sort(key=lambda t: int(batches[t[0]].ids[t[1]]))
        # C-ABI arguments of every call, built once
        self._calls = []
        for _, _, dg, out, W in self.graphs:
            g = dg.c_struct()
            self._calls.append((g, ctypes.byref(g), out.data_ptr() if dg.m else None,
                                None if W is None else W.data_ptr()))

    def step(self, stream=None):
        lib = load_jaccard_lib()
        st = api._stream(stream)
        for g, gref, out, W in self._calls:
            if W is not None:
                code = lib.jc_vertex_weights(gref, W, st)
                if code != 0:
                    raise api.JaccardError(code, "jc_vertex_weights")
            code = lib.jc_jaccard(gref, W, out, api.EDGE_CENTRIC, st)
            if code != 0:
                raise api.JaccardError(code, "jc_jaccard")
