"""Benchmark: Jaccard coefficients/s on one MI355X.

Default workload (the configuration the headline metric is quoted on):
RMAT scale 22, edge factor 16, a/b/c = .57/.19/.19, symmetrised, deduped,
sorted int32 CSR (~1.28e8 stored edges), float32 edge weights uniform in
(0,1]. One step = vertex weights from edge weights (the reference driver's
weighted preamble) + the edge-centric weighted Jaccard of every stored edge.
Inputs are resident in HBM before the timed region.

Prints one JSON line. ``roofline`` uses the algorithmic bytes
  B = 4*sum_{stored (u,v)} min(d_u,d_v) + 4(n+1) + 4m + 4m  (+ 4m + 8n weighted)
over the dominant kernel's average time per call (HIP events on its stream
around its two launches, the hash-table and the bitmap task lists of
``k_owner_block``).
``cpu_baseline`` times the CPU restatement of the reference's edge-centric
loop (binary search per shorter-row element) on a bounded random sample of
the same graph's edges, with all host threads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sycl-jaccard_amd"))

METRIC = "Jaccard coeffs/s + % HBM roofline; batched graphs/s & p99 latency vs plain"
HBM_PEAK = 8.0e12

WORKLOADS = {
    # name: (generator kwargs, weighted, description)
    "rmat22_weighted": (dict(kind="rmat", scale=22, edge_factor=16, seed=1), True,
                        "RMAT-22 ef16 weighted (0,1]"),
    "rmat16": (dict(kind="rmat", scale=16, edge_factor=16, seed=1), False, "RMAT-16 ef16 unweighted"),
    "chunglu20": (dict(kind="chunglu", n=1 << 20, seed=1), False,
                  "Chung-Lu 2^20 gamma 2.1 mean 32 max 2^17 unweighted"),
}


# rocprofv3 summary of the dominant kernel's HBM traffic per workload: bytes
# from the TCC read/write request counts by request size (see the file for
# the commands and the copy-kernel check).
TRAFFIC_PROFILES = {"rmat22_weighted": "profiles/owner_block_rmat22_weighted_pmc.md",
                    "chunglu20": "profiles/owner_block_chunglu20_pmc.md"}


def profile_value(workload, key, cast=float):
    """A `key = value` line of the committed counter summary of the workload, or None."""
    rel = TRAFFIC_PROFILES.get(workload)
    if rel is None or not os.path.exists(os.path.join(ROOT, rel)):
        return None
    for line in open(os.path.join(ROOT, rel)):
        if line.startswith(key + " ="):
            return cast(line.split("=")[1])
    return None


def measured_traffic(workload):
    """HBM bytes per k_owner_block launch from the committed counter summary, or None."""
    return profile_value(workload, "traffic_bytes_per_launch", int)


def make_graph(spec, weighted):
    from jaccard_amd import graphs
    kind = spec["kind"]
    if kind == "rmat":
        return graphs.rmat(spec["scale"], spec["edge_factor"], seed=spec["seed"], weighted=weighted)
    return graphs.chung_lu(n=spec["n"], seed=spec["seed"], weighted=weighted)


def compulsory_bytes(g, weighted):
    """CSR (and weights) read once plus the output written once."""
    b = 4 * (g.n + 1) + 4 * g.m + 4 * g.m
    if weighted:
        b += 4 * g.m + 8 * g.n
    return int(b)


def algorithmic_bytes(g, st, weighted):
    return int(4 * st["sum_min_deg"]) + compulsory_bytes(g, weighted)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(g, budget_s=15.0, seed=0, runs=3):
    """Reference EC loop (oracle restatement) on a random edge sample, all host
    threads: a sample sized to about budget_s / runs seconds, timed `runs`
    times, median rate."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = oracle.max_threads()
    rng = np.random.default_rng(seed)
    # size the sample: double until one pass takes a tenth of a run's share
    n_s, dt = 2000, 0.0
    while True:
        edges = rng.integers(0, g.m, n_s)
        t0 = time.perf_counter()
        oracle.ec_edges(g, edges, threads)
        dt = time.perf_counter() - t0
        if dt >= 0.1 * budget_s / runs or n_s >= g.m:
            break
        n_s *= 2
    n_s = int(min(g.m, max(n_s, n_s * (budget_s / runs) / max(dt, 1e-9))))
    edges = rng.integers(0, g.m, n_s)
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.ec_edges(g, edges, threads)
        rates.append(n_s / (time.perf_counter() - t0))
    rates.sort()
    med = rates[len(rates) // 2]
    return {"value": med, "unit": "coeffs/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "runs": [round(r, 1) for r in rates],
            "sample": f"{n_s} uniformly sampled stored edges of the same graph, reference edge-centric "
                      f"binary-search loop (jaccard.cpp:711-742) restated in C/OpenMP, median of {runs} runs "
                      f"(about {n_s / med:.1f} s each)"}


def oracle_check(g, out, W, weighted, n_sample=200000, seed=3):
    """Part of the CPU-baseline leg: the oracle, as the checker, on a random
    sample of stored edges of the measured step's device output. Weighted:
    the FP64 quotient (worst relative error reported against the 1e-6
    tolerance); unweighted: the reference loop's fp32 result, bit for bit."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rng = np.random.default_rng(seed)
    k = np.arange(g.m) if n_sample >= g.m else np.unique(rng.integers(0, g.m, n_sample))
    got = out.cpu().numpy()[k]
    if weighted:
        rows = (np.searchsorted(g.offsets, k, side="right") - 1).astype(np.int32)
        cols = g.indices[k].astype(np.int32)
        ref = oracle.pairs_weighted_f64(g, W.cpu().numpy().astype(np.float64), rows, cols)
        nz = ref != 0
        rel = np.abs(got[nz].astype(np.float64) - ref[nz]) / np.abs(ref[nz])
        zero_ok = bool(np.all(got[~nz] == 0))
        return {"sample_edges": int(k.size), "max_rel_err": float(rel.max()) if rel.size else 0.0,
                "tolerance": 1e-6, "zeros_exact": zero_ok, "reference": "FP64 oracle, per-pair weighted quotient"}
    ref = oracle.ec_edges(g, k)
    return {"sample_edges": int(k.size), "bit_exact": bool(np.array_equal(got.view(np.uint32), ref.view(np.uint32))),
            "reference": "oracle restatement of the reference edge-centric loop (jaccard.cpp:711-742)"}


def run_kernel_level(dg, W, out):
    """The reference's own kernel sequences through the kernel-level entry
    points, on the bench graph: EC as Jaccard_ec_scan + Jaccard_ec_unweighted
    (jaccard.cpp:654-782), VC as RowSum + Fill + Is + Jw (jaccard.cpp:858-896,
    fp32 buffers as the reference keeps them). Median of 3 synchronized runs;
    the EC result is compared bit for bit with the fused unweighted call."""
    import torch
    from jaccard_amd import api
    m, n, dev = dg.m, dg.n, dg.offsets.device
    dest = torch.empty(m, dtype=torch.int32, device=dev)
    wj = torch.empty(m, dtype=torch.float32, device=dev)
    work = torch.empty(n, dtype=torch.float32, device=dev)
    wi = torch.empty(m, dtype=torch.float32, device=dev)
    ws = torch.empty(m, dtype=torch.float32, device=dev)

    def ec():
        api.ec_scan(dg, dest, wj)
        api.ec_unweighted(dg, dest, wj)

    def vc():
        api.rowsum(dg, work, v=W)
        api.fill(wi, 0.0)
        api.intersect(dg, work, wi, ws, v=W)
        api.jw(wi, ws, wj)

    res = {}
    for name, fn in (("ec_scan_plus_ec_unweighted_ms", ec), ("rowsum_fill_is_jw_ms", vc)):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        res[name] = sorted(ts)[1]
    ec()
    fused = api.jaccard(dg, torch.empty_like(out), edge_centric=True)
    torch.cuda.synchronize()
    res["ec_kernel_level_equals_fused"] = bool(torch.equal(wj.view(torch.int32), fused.view(torch.int32)))
    del dest, wj, work, wi, ws, fused
    return res


def run_batched(args):
    """Batched workload: 4096 RMAT graphs (n log-uniform in [2^8, 2^16], edge
    factor uniform in [4, 32], half weighted) in one captured HIP graph vs one
    plain call per graph."""
    import numpy as np
    import torch
    from jaccard_amd import batch
    t0 = time.time()
    bu, bw = batch.make_batches(count=args.batch_count, seed=1)
    print(f"[bench] batch of {bu.count + bw.count} graphs, m={bu.union.m + bw.union.m}, "
          f"gen {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    br = batch.BatchRunner([bu, bw], dev)
    pr = batch.PlainRunner([bu, bw], dev)
    # plain: 4096 drop-in calls back to back on one stream
    pr.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.batch_reps):
        pr.step()
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t1) / args.batch_reps
    print(f"[bench] plain {plain_s * 1e3:.1f} ms per batch", file=sys.stderr, flush=True)
    # device work of the plain calls (kernel time, HIP events per launch; one
    # stream, so they do not overlap): the part of the plain time that is
    # compute, against the rest (launches, allocations, host work)
    from jaccard_amd import api
    api.profile_reset()
    api.profile_enable(True)
    pr.step()
    torch.cuda.synchronize()
    api.profile_enable(False)
    _, plain_dev_ms = api.profile_query(None)
    api.profile_reset()
    tune = br.autotune(reps=2)
    br.capture()
    br.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.batch_reps):
        br.replay()
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t1) / args.batch_reps
    print(f"[bench] batched {batched_s * 1e3:.1f} ms per batch, tuning {br.tuning}", file=sys.stderr, flush=True)
    same = all(torch.equal(br.member_output(k, i).view(torch.int32), out.view(torch.int32))
               for k, i, _, out, _ in pr.graphs)
    ngraphs = bu.count + bw.count
    return {"graphs": ngraphs, "stored_edges": int(bu.union.m + bw.union.m),
            "batched_graphs_per_s": ngraphs / batched_s, "plain_graphs_per_s": ngraphs / plain_s,
            "speedup": plain_s / batched_s, "batched_ms": batched_s * 1e3, "plain_ms": plain_s * 1e3,
            "identical_outputs": bool(same), "tuning": [list(t) for t in br.tuning],
            "plain_device_ms": plain_dev_ms,
            "coeffs_per_s_batched": (bu.union.m + bw.union.m) / batched_s}


def profile_batched(args):
    """The batched form alone: generation, autotuning and capture, then
    args.steps timed replays of the captured step (no plain calls)."""
    import torch
    from jaccard_amd import batch
    bu, bw = batch.make_batches(count=args.batch_count, seed=1)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    br = batch.BatchRunner([bu, bw], dev)
    br.autotune(reps=1)
    br.capture()
    br.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        br.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t1) * 1e3 / args.steps
    print(f"[bench] batched only: {ms:.1f} ms per replay, tuning {br.tuning}", file=sys.stderr, flush=True)
    return {"mode": "profile_batched", "graphs": bu.count + bw.count, "stored_edges": int(bu.union.m + bw.union.m),
            "batched_ms": ms, "replays": args.steps, "tuning": [list(t) for t in br.tuning]}


def run_served(args):
    """Serving workload: 64 resident RMAT graphs (scales 14..20 cycling, edge
    factor 16, float32 weights in (0,1]); requests pick a graph uniformly and are
    whole-graph weighted Jaccard (probability 1/2) or k vertex pairs (k
    log-uniform in [1, 65536], half stored edges, half uniform pairs).
    Continuous batching vs one plain launch per request (FIFO), first as a burst
    (all requests at t=0: throughput), then as a Poisson stream at a fixed QPS
    below the plain path's capacity (latency)."""
    import numpy as np
    import torch
    from jaccard_amd import serve
    t0 = time.time()
    gs = serve.serving_graphs(count=64, seed=7)
    dev = torch.device("cuda:0")
    dg = serve.upload(gs, dev, weighted=True)
    print(f"[bench] serving: 64 graphs, m={sum(g.m for g in gs)}, setup {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    cont = serve.Server(gs, dev, mode=serve.CONTINUOUS, slots=4, hbm_budget_bytes=1 << 30, device_graphs=dg)
    fifo = serve.Server(gs, dev, mode=serve.FIFO, device_graphs=dg)
    out = {"graphs": 64, "stored_edges": int(sum(g.m for g in gs)), "arena_bytes": cont.arena_bytes(),
           "hbm_budget_bytes": 1 << 30, "slots": 4}

    def both(reqs, tag):
        # one result buffer per request, so the identity check covers every request
        res_c = serve.result_buffers(gs, reqs, dev)
        res_f = serve.result_buffers(gs, reqs, dev)
        cont.run(reqs[: min(8, len(reqs))], res_c[: min(8, len(reqs))])  # warm
        st_c, info_c = cont.run(reqs, res_c)
        st_f, info_f = fifo.run(reqs, res_f)
        assert (st_c["status"] == 0).all() and (st_f["status"] == 0).all()
        torch.cuda.synchronize()
        same = all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip(res_c, res_f))
        del res_c, res_f
        c, f = serve.latency_summary(reqs, st_c), serve.latency_summary(reqs, st_f)
        c.update(batches=info_c["batches"], max_in_flight=info_c["max_in_flight"],
                 peak_budget_bytes=info_c["peak_budget_bytes"])
        print(f"[bench] serving {tag}: continuous {c['requests_per_s']:.1f} req/s p50 {c['p50_ms']:.2f} "
              f"p99 {c['p99_ms']:.2f} ms | plain {f['requests_per_s']:.1f} req/s p50 {f['p50_ms']:.2f} "
              f"p99 {f['p99_ms']:.2f} ms | identical {same}", file=sys.stderr, flush=True)
        return {"continuous": c, "plain": f, "identical_outputs": bool(same),
                "throughput_ratio": c["requests_per_s"] / f["requests_per_s"]}

    burst = serve.request_stream(gs, args.serve_requests, qps=0.0, seed=11)
    out["burst"] = both(burst, "burst")
    # the burst's device work: kernel time of the plain path (one request at a
    # time, so the per-kernel HIP events do not overlap); no schedule can
    # finish the burst faster than this, which bounds the continuous/plain ratio
    from jaccard_amd import api

    def plain_device_ms(reqs):
        res = serve.result_buffers(gs, reqs, dev)
        torch.cuda.synchronize()
        api.profile_reset()
        api.profile_enable(True)
        fifo.run(reqs, res)
        api.profile_enable(False)
        _, ms = api.profile_query(None)
        owner = sum(api.profile_query(k)[1] for k in ("owner_block", "owner_small"))
        api.profile_reset()
        del res
        return ms, owner

    dev_ms, _ = plain_device_ms(burst)
    # the burst's full-graph requests: the owner-engine passes no schedule can
    # avoid (continuous batching merges them into fewer launches but still
    # streams every requested row once) bound the continuous throughput
    full = [q for q in burst if q.kind == serve.REQ_FULL]
    full_ms, full_owner_ms = plain_device_ms(full)
    b = out["burst"]
    b["device_ms"] = dev_ms
    b["bound_throughput_ratio"] = b["plain"]["wall_s"] * 1e3 / dev_ms if dev_ms else None
    b["full_requests"] = len(full)
    b["full_device_ms"] = full_ms
    b["full_owner_kernel_ms"] = full_owner_ms
    b["owner_bound_throughput_ratio"] = b["plain"]["wall_s"] * 1e3 / full_owner_ms if full_owner_ms else None
    print(f"[bench] serving burst device work {dev_ms:.1f} ms (full-graph requests {full_ms:.1f} ms, of which "
          f"owner kernels {full_owner_ms:.1f} ms): continuous/plain at most "
          f"{b['owner_bound_throughput_ratio']:.2f}x", file=sys.stderr, flush=True)
    plain_cap = out["burst"]["plain"]["requests_per_s"]
    # below the plain path's capacity (latency at moderate load), then above it
    # (the plain queue grows without bound; continuous batching absorbs it)
    for key, qps, seed in (("poisson", args.serve_qps or 0.7 * plain_cap, 12),
                           ("poisson_overload", 1.15 * plain_cap, 13)):
        stream = serve.request_stream(gs, args.serve_requests, qps=qps, seed=seed)
        out[key] = both(stream, f"{key} {qps:.0f} qps")
        out[key]["qps"] = qps
        out[key]["qps_over_plain_capacity"] = qps / plain_cap
    cont.close()
    fifo.close()
    return out


def tuning_sweep(args):
    """Every jc_tuning field over graph families (the data behind
    jc_config_from_stats): one JSON line per (graph, configuration) with the
    median call time of `reps` synchronized jc_jaccard_tuned calls and
    whether the result equals the first configuration's bit for bit."""
    import itertools
    import numpy as np
    import torch
    from jaccard_amd import api, batch, graphs
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    fields = ("small_max", "small_blocks_per_cu", "bitmap_blocks_per_cu", "hash_table")
    full = list(itertools.product((128, 96, 64, 32), (8, 16, 32), (1, 2), (0, 1)))
    small = list(itertools.product((128, 64), (16, 32), (1, 2), (0, 1)))
    bb = [(128, 32, 1, 0), (128, 32, 2, 0)]

    def run(name, g, weighted, grid, reps=3):
        dg = api.DeviceCSR.from_host(g, dev)
        W = api.vertex_weights(dg) if weighted else None
        out = torch.empty(g.m, dtype=torch.float32, device=dev)
        ws = torch.empty(api.workspace_size(g.n, g.m, weighted, True), dtype=torch.uint8, device=dev)
        st = api.degree_stats(dg)
        chosen = api.config_from_stats(st, g.n, weighted)[1]
        print(f"[bench] sweep {name}: n={g.n} m={g.m} chosen {chosen}", file=sys.stderr, flush=True)
        ref = None
        for cand in grid:
            t = dict(zip(fields, cand))
            call = lambda: api.jaccard(dg, out, weights=W, workspace=ws, flags=api.FLAG_ASSUME_SYMMETRIC, tuning=t)
            call()
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                call()
                torch.cuda.synchronize()
                ts.append((time.perf_counter() - t0) * 1e3)
            bits = out.view(torch.int32).clone()
            same = True if ref is None else bool(torch.equal(bits, ref))
            ref = bits if ref is None else ref
            print(json.dumps({"graph": name, "weighted": weighted, "n": g.n, "m": g.m, "tuning": t,
                              "chosen": all(t[k] == chosen[k] for k in fields),
                              "ms": sorted(ts)[len(ts) // 2], "same_bits": same}), flush=True)
        del dg, W, out, ws
        torch.cuda.empty_cache()

    run("rmat16", graphs.rmat(16, 16, seed=1), False, full)
    run("rmat18w", graphs.rmat(18, 16, seed=1, weighted=True), True, full)
    run("rmat20", graphs.rmat(20, 16, seed=1), False, full)
    run("chunglu20", graphs.chung_lu(1 << 20, seed=1), False, full)
    run("rmat22w", graphs.rmat(22, 16, seed=1, weighted=True), True, small, reps=2)
    for sc in (14, 16, 17, 19, 20):
        run(f"rmat{sc}w", graphs.rmat(sc, 16, seed=1, weighted=True), True, bb)
    run("chunglu18w", graphs.chung_lu(1 << 18, seed=1, weighted=True), True, bb)
    bu, bw = batch.make_batches(count=args.batch_count, seed=1)
    run("batch_unweighted", bu.union, False, small)
    run("batch_weighted", bw.union, True, small)


def run_ingest(args):
    """Graph ingest of the RMAT-22 ef16 draws (6.7e7 directed edges):
    jc_csr_from_edges on the device (inputs resident) vs the host C++/OpenMP
    builder, same CSR."""
    import ctypes
    import numpy as np
    import torch
    from jaccard_amd import api, graphs
    from jaccard_amd._native import load_graph_lib
    lib = load_graph_lib()
    scale, ef = 22, 16
    e = ef << scale
    s = np.empty(e, np.int32)
    d = np.empty(e, np.int32)
    lib.jcg_rmat_edges(scale, 0, float(ef), graphs.RMAT_A, graphs.RMAT_B, graphs.RMAT_C, 1, s.ctypes.data,
                       d.ctypes.data, e)
    dev = torch.device("cuda:0")
    ds, dd = torch.from_numpy(s).to(dev), torch.from_numpy(d).to(dev)
    api.csr_from_edges(1 << scale, ds, dd)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = api.csr_from_edges(1 << scale, ds, dd)
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    h = graphs.from_edges(1 << scale, s, d)
    t_host = time.perf_counter() - t0
    same = bool(np.array_equal(g.indices.cpu().numpy(), h.indices) and np.array_equal(g.offsets.cpu().numpy(), h.offsets))
    print(f"[bench] ingest: device {t_dev * 1e3:.1f} ms, host {t_host * 1e3:.1f} ms, same {same}", file=sys.stderr,
          flush=True)
    return {"edges_in": e, "stored_edges": int(g.m), "device_ms": t_dev * 1e3, "host_ms": t_host * 1e3,
            "host_threads": os.cpu_count(), "identical_csr": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22_weighted", choices=sorted(WORKLOADS))
    ap.add_argument("--algo", default="ec", choices=["ec", "vc"])
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--extras", action=argparse.BooleanOptionalAction, default=True,
                    help="also time VC (the vertex-centric pipeline) and unweighted EC (default on)")
    ap.add_argument("--batched", action=argparse.BooleanOptionalAction, default=True,
                    help="also run the batched workload, 4096 graphs (default on)")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--served", action=argparse.BooleanOptionalAction, default=True,
                    help="also run the serving workload, 64 graphs (default on)")
    ap.add_argument("--hash-table32", action="store_true",
                    help="32-bit hash tables for every hash-table task (jc_tuning.hash_table = 1)")
    ap.add_argument("--core-only", action="store_true",
                    help="only the headline step (no extras, batched or served; for profiling)")
    ap.add_argument("--ingest", action="store_true", help="also time device graph ingest of the RMAT-22 draws")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-qps", type=float, default=0.0, help="Poisson rate (0: 0.7 x plain capacity)")
    ap.add_argument("--profile-batched", action="store_true",
                    help="only the batched form's captured step, --steps replays after autotuning "
                         "(for a kernel trace of it alone); prints one JSON line about it")
    ap.add_argument("--tuning-sweep", action="store_true",
                    help="only time every jc_tuning field over graph families (JSON line each)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="also run 3 device-to-device copies of BYTES (a known-size kernel that "
                         "calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.core_only:
        args.extras = args.batched = args.served = False
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from jaccard_amd import api

    if args.profile_batched:
        print(json.dumps(profile_batched(args)))
        return
    if args.tuning_sweep:
        tuning_sweep(args)
        return

    spec, weighted, desc = WORKLOADS[args.workload]
    t0 = time.time()
    g = make_graph(spec, weighted)
    st = g.stats()
    print(f"[bench] {desc}: n={g.n} m={g.m} sum_min={st['sum_min_deg']} max_deg={st['max_deg']} "
          f"gen {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    dg = api.DeviceCSR.from_host(g, dev)
    out = torch.empty(g.m, dtype=torch.float32, device=dev)
    W = torch.empty(g.n, dtype=torch.float32, device=dev)
    ec = args.algo == "ec"
    stream = torch.cuda.current_stream()

    tuning = {"hash_table": 1} if args.hash_table32 else None

    def step():
        if weighted:
            api.vertex_weights(dg, out=W, stream=stream)
            api.jaccard(dg, out, weights=W, edge_centric=ec, stream=stream, tuning=tuning)
        else:
            api.jaccard(dg, out, edge_centric=ec, stream=stream, tuning=tuning)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    api.profile_reset()
    api.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    api.profile_enable(False)
    kern = {name: api.profile_query(name) for name in api.profile_names()}
    api.profile_reset()
    ms_per_step = wall * 1e3 / args.steps
    value = g.m / (wall / args.steps)
    nb, tb = kern.get("owner_block", (0, 0.0))
    dev_ms = sum(t for _, t in kern.values()) / args.steps
    B = algorithmic_bytes(g, st, weighted)
    achieved = B / (tb / nb * 1e-3) if nb else 0.0
    print(f"[bench] {ms_per_step:.2f} ms/step, device {dev_ms:.2f} ms/step, owner_block "
          f"{(tb / nb if nb else 0):.2f} ms", file=sys.stderr, flush=True)

    # one C-ABI call at a time, host clock around call + synchronize
    calls = []
    for _ in range(3):
        t1 = time.perf_counter()
        api.jaccard(dg, out, weights=W if weighted else None, edge_centric=ec, stream=stream, tuning=tuning)
        torch.cuda.synchronize()
        calls.append((time.perf_counter() - t1) * 1e3)
    # the committed counter summary describes the default (quotient-table) kernel
    traffic = None if args.hash_table32 else measured_traffic(args.workload)
    ob_ms = tb / nb if nb else None
    Bc = compulsory_bytes(g, weighted)
    extras = {"device_ms_per_step": dev_ms,
              "kernel_ms_per_step": {k: t / args.steps for k, (c, t) in kern.items()},
              "step_roofline_frac": B / (dev_ms * 1e-3) / HBM_PEAK if dev_ms else None,
              "call_ms": sorted(calls)[1],
              "compulsory_bytes": Bc,
              "compulsory_frac": Bc / (dev_ms * 1e-3) / HBM_PEAK if dev_ms else None,
              "tcc_hit_rate": profile_value(args.workload, "tcc_hit_rate"),
              "sum_min_deg": st["sum_min_deg"], "max_deg": st["max_deg"]}
    if args.extras:
        for name, kw in (("vc", dict(edge_centric=False)), ("ec_unweighted", dict(edge_centric=True))):
            wts = W if (weighted and name == "vc") else None
            api.jaccard(dg, torch.empty_like(out), weights=wts, **kw)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(3):
                api.jaccard(dg, torch.empty_like(out), weights=wts, **kw)
            torch.cuda.synchronize()
            extras[f"{name}_ms"] = (time.perf_counter() - t1) * 1e3 / 3
        extras["kernel_level"] = run_kernel_level(dg, W if weighted else None, out)

    if args.batched:
        extras["batched"] = run_batched(args)

    if args.served:
        extras["served"] = run_served(args)

    if args.ingest:
        extras["ingest"] = run_ingest(args)

    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        extras["calibration_copy_bytes"] = args.calibrate_copy
        del src, dst

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(g, args.cpu_budget)
        extras["oracle_check"] = oracle_check(g, out, W if weighted else None, weighted)
        print(f"[bench] oracle check: {extras['oracle_check']}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC, "value": value, "unit": "coeffs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if weighted else "int32", "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "algorithm": "edge_centric" if ec else "vertex_centric",
                   "n": g.n, "m": g.m, "weighted": weighted, **{k: v for k, v in spec.items() if k != "kind"},
                   "hash_tables": "32-bit" if args.hash_table32 else "quotient below 2^24 vertices"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": "k_owner_block",
                     "algorithmic_bytes": B,
                     # the counted bytes (profiles/) over the same measured kernel time
                     "physical_achieved": traffic / (ob_ms * 1e-3) / 1e9 if (traffic and ob_ms) else None,
                     "physical_frac": traffic / (ob_ms * 1e-3) / HBM_PEAK if (traffic and ob_ms) else None},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
