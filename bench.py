#!/usr/bin/env python3
"""Batched vs per-source multi-source BFS on one MI355X.

A step is one pass of the batched bit-parallel engine over the whole source
batch of the workload (default: the LiveJournal-shaped Chung-Lu graph,
n = 4.85M, m ~ 69M, 1,024 degree-weighted sources), with the graph and the
S x n level/parent output tables resident in HBM.  The plain per-source engine
(one top-down push launch sequence per source) runs the same batch once, and
its tables must match the batched ones bit for bit.  The CPU baseline times
the multi-threaded CPU restatements of the same path (bit-parallel and
per-source, both with the |E|/20 push/pull switch and the same output rule) on
bounded samples of the same sources, on every usable host core and on one.
The line also carries a serving sub-object: the web-crawl-shaped graph
resident on the device, queries at 2,000/s offered, and the knee.

Prints one JSON line.  GTEPS = sum over sources of the out-degrees of the
vertices each source reaches, divided by wall time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "xblossom-sigmetrics27-ae_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "GTEPS batched vs per-source on 1 MI355X (bit-exact); %HBM roofline; p99 query ms"
HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")

# rocprofv3 kernel names of the engine's kernel classes (for the PMC traffic lookup).
KERNEL_SYMBOL = {
    "emit": "k_bp_emit_light",
    "pull": "k_bp_pull_fused",
    "pull_hub": "k_bp_pull_hub",
    "push": "k_bp_push",
    "plain_push": "k_plain_push",
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name: str):
    import xbmsbfs as xb

    if name == "livejournal":
        g, src = xb.workload_livejournal_like()
        desc = "LiveJournal-shaped Chung-Lu (gamma 2.1, seed 4), 1,024 degree-weighted sources"
    elif name == "amazon":
        g, src = xb.workload_amazon_like()
        desc = "Amazon-shaped Chung-Lu (gamma 2.3, seed 3), 256 uniform sources"
    elif name == "rmat16":
        g, src = xb.workload_rmat_s16()
        desc = "R-MAT scale 16 edge factor 16 (seed 1), 64 uniform sources (seed 2)"
    else:
        raise SystemExit(f"unknown workload {name}")
    return g, src, desc


# Engine sources whose hash keys a committed PMC summary: a summary measured
# on other kernels, or on another graph, is not used.
ENGINE_SOURCES = ("xblossom-sigmetrics27-ae_amd/csrc/xb_kernels.hip", "xblossom-sigmetrics27-ae_amd/csrc/xb_host.cpp",
                  "xblossom-sigmetrics27-ae_amd/csrc/xb_internal.hpp")
TRAFFIC_SOURCE = ("rocprofv3 PMC, L2-to-fabric requests by size (TCC_EA0_RDREQ_32B/64B/128B x 32/64/128 B read, "
                  "TCC_EA0_WRREQ_64B x 64 B + other writes x 32 B), one pass per counter set; a 1 GiB copy reads "
                  "exactly 1.00x its size with these counters (FETCH_SIZE books 128-B requests as 64 B: 0.50x). "
                  "Infinity Cache hits are included, so this is an upper bound on HBM bytes")


def engine_source_hash() -> str:
    import hashlib

    h = hashlib.sha256()
    for rel in ENGINE_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def profile_values(workload: str, kernel: str, graph_checksum: str = None) -> dict:
    """The key=value lines of a committed rocprofv3 PMC summary,
    profiles/pmc_<kernel>_<workload>.txt, when it was measured on the engine
    sources of this run (source_hash) and, if given, on the same graph
    (graph_checksum); otherwise {} with the reason under "_stale".  Keys:
    traffic_bytes_per_launch (read + write bytes per dispatch of `kernel`),
    step_traffic_bytes (every engine kernel and fill of one step), and the
    configuration it was measured in (profiled_engine,
    profiled_words_per_tile)."""
    sym = KERNEL_SYMBOL.get(kernel, kernel)
    path = os.path.join(PROFILES, f"pmc_{sym}_{workload}.txt")
    vals = {}
    if not os.path.exists(path):
        return {"_stale": f"no {os.path.basename(path)}"}
    with open(path) as f:
        for line in f:
            key, eq, val = line.strip().partition("=")
            if eq and key.isidentifier():
                vals[key] = val
    if vals.get("source_hash") != engine_source_hash():
        return {"_stale": "profiled on other engine sources"}
    if graph_checksum and vals.get("graph_checksum") not in (None, graph_checksum):
        return {"_stale": "profiled on another graph"}
    return vals


def traffic_from_profiles(workload: str, kernel: str, graph_checksum: str = None):
    """Measured bytes per launch of `kernel` from the committed PMC summary, or None."""
    v = profile_values(workload, kernel, graph_checksum).get("traffic_bytes_per_launch")
    return float(v) if v else None


def host_cores():
    """(physical cores, this process's CPU share, threads used).  Physical
    cores follow the reference's policy, cores per socket x sockets
    (scripts/4-expr/common.sh:164-184), counted here as distinct (physical
    id, core id) pairs of /proc/cpuinfo.  The threads used are capped by the
    CPUs this process may run on and by OMP_NUM_THREADS when it is set."""
    pairs, pid = set(), None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "physical id":
                    pid = v.strip()
                elif k == "core id":
                    pairs.add((pid, v.strip()))
    except OSError:
        pass
    physical = len(pairs) or (os.cpu_count() or 1)
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    used = min(physical, share)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        used = min(used, int(omp))
    return physical, share, max(1, used)


def trimmed_mean(xs):
    """Mean with the minimum and maximum dropped (the reference's timing
    rule, apps/xb/src/include/common/timing_guard.h:43-71)."""
    xs = sorted(xs)
    if len(xs) > 2:
        xs = xs[1:-1]
    return sum(xs) / len(xs)


def cpu_baseline(g, src, rounds: int):
    """Times the multi-threaded CPU restatements (oracle/msbfs_cpu_baseline.c)
    on bounded samples of the workload's sources: the bit-parallel mode (64
    sources per word, levels parallel over vertices in chunks of 64) and the
    per-source direction-optimizing mode (threads over sources), each with
    every usable core and with one thread, `rounds` rounds each on different
    sources, GTEPS as the mean with the minimum and maximum dropped."""
    import oracle

    physical, share, threads = host_cores()
    log(f"cpu baseline: {physical} physical cores (cores/socket x sockets), {share} CPUs in this process's "
        f"share, {threads} threads used")
    plan = {"bitpar": ((threads, 64), (1, 64)), "dir": ((threads, threads), (1, 1))}
    modes = {}
    total_s = 0.0
    for mode, cfgs in plan.items():
        res = {}
        for t, S in cfgs:
            vals = []
            for r in range(rounds):
                b0 = (r * S) % max(1, src.size - S + 1)
                part = src[b0:b0 + S]
                t0 = time.perf_counter()
                _, _, e = oracle.msbfs_cpu(g.row_offsets, g.col_indices, part, threads=t, mode=mode)
                dt = time.perf_counter() - t0
                total_s += dt
                vals.append(e / dt / 1e9)
            res["all_cores" if t == threads else "one_thread"] = {
                "GTEPS": round(trimmed_mean(vals), 4), "threads": t, "sources_per_round": int(S),
                "rounds_GTEPS": [round(v, 4) for v in vals]}
        modes[mode] = res
    best = max(("bitpar", "dir"), key=lambda m: modes[m]["all_cores"]["GTEPS"])
    return {
        "value": modes[best]["all_cores"]["GTEPS"],
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "mode": best,
        "host_physical_cores": physical,
        "cpu_share": share,
        "rounds": rounds,
        "statistic": "mean of the rounds with the minimum and maximum dropped",
        "modes": modes,
        "sample": (f"{rounds} rounds per mode and thread count on sources of the same workload: bit-parallel "
                   f"64 sources per round, per-source {threads} (all cores) or 1 (one thread) per round; "
                   f"levels, lowest-id parents and the |E|/20 push/pull switch as the engine; "
                   f"{total_s:.1f} s in all"),
    }


def bench_uneven(args) -> int:  # noqa: C901
    """The uneven batch: 32 R-MAT graphs (scales 12-22) with 1-512 sources each.

    Batched = one replay of the captured batch: the persistent kernel (every
    graph's traversal) and the emit launches of the one-tile teams after it;
    reported beside the host-driven single-graph engine run graph by graph and
    the plain per-source engine on every source."""
    import numpy as np
    import torch

    import xbmsbfs as xb

    t0 = time.perf_counter()
    work = xb.workload_uneven_batch()
    log(f"uneven batch: {len(work)} graphs, {sum(g.m for g, _ in work)} edges, "
        f"{sum(s.size for _, s in work)} sources ({time.perf_counter() - t0:.1f}s)")
    hgs = [xb.HGraph(g.row_offsets, g.col_indices) for g, _ in work]
    srcs = [s for _, s in work]
    lv = [torch.empty((s.size, g.nnodes), dtype=torch.int32, device="cuda") for s, g in zip(srcs, hgs)]
    pa = [torch.empty((s.size, g.nnodes), dtype=torch.int32, device="cuda") for s, g in zip(srcs, hgs)]
    tun = {"words_per_tile": args.words_per_tile} if args.words_per_tile else {}
    batch = xb.Batch(hgs, srcs, [t.data_ptr() for t in lv], [t.data_ptr() for t in pa], **tun)
    for _ in range(args.warmup):
        batch.run()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    ms_graph = []
    for _ in range(args.steps):
        ms_graph.append(batch.run().ms_total)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t_start) / args.steps
    edges = sum(xb.XBlossomGpuEngine10(h).CountTraversed(l.data_ptr(), s.size) for h, l, s in zip(hgs, lv, srcs))
    gteps = edges / t_batch / 1e9
    log(f"persistent batch: {t_batch * 1e3:.2f} ms, {gteps:.1f} GTEPS")
    sums = [(l.sum().item(), p.sum().item()) for l, p in zip(lv, pa)]
    # Host-driven engine, graph by graph, into the same tables.
    engines = [xb.XBlossomGpuEngine10(h) for h in hgs]
    for e, s, l, p in zip(engines, srcs, lv, pa):
        e.MultiSourceBFS(s, l.data_ptr(), p.data_ptr(), mode=xb.XB_MODE_BATCHED, device_out=True, **tun)
    torch.cuda.synchronize()
    th0 = time.perf_counter()
    for e, s, l, p in zip(engines, srcs, lv, pa):
        e.MultiSourceBFS(s, l.data_ptr(), p.data_ptr(), mode=xb.XB_MODE_BATCHED, device_out=True, **tun)
    torch.cuda.synchronize()
    t_host = time.perf_counter() - th0
    exact = sums == [(l.sum().item(), p.sum().item()) for l, p in zip(lv, pa)]
    log(f"host-driven per graph: {t_host * 1e3:.2f} ms, equal sums: {exact}")
    plain = None
    if not args.no_plain:
        tp0 = time.perf_counter()
        for e, s, l, p in zip(engines, srcs, lv, pa):
            e.MultiSourceBFS(s, l.data_ptr(), p.data_ptr(), mode=xb.XB_MODE_PLAIN, device_out=True)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - tp0
        plain = {"gteps": round(edges / t_plain / 1e9, 3), "ms_per_batch": round(t_plain * 1e3, 2)}
        exact = exact and sums == [(l.sum().item(), p.sum().item()) for l, p in zip(lv, pa)]
        log(f"plain: {t_plain * 1e3:.1f} ms")
    st = batch.last_stats
    achieved = st.kernel_bytes[11] / (st.kernel_ms[11] * 1e-3) / 1e9
    avg_launch_ms = float(np.mean(ms_graph))
    traffic = traffic_from_profiles("uneven", "k_bp_persistent")
    line = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batch * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "uneven batch: 32 R-MAT graphs, scales 12-22, edge factor 8-32 (seed 5), "
                               "1-512 sources each (seed 6)",
                   "graphs": len(work), "m_total": int(sum(g.m for g, _ in work)),
                   "sources": int(sum(s.size for s in srcs)), "edges_traversed_per_step": int(edges),
                   "engine": "persistent"},
        "host_driven_per_graph": {"gteps": round(edges / t_host / 1e9, 3), "ms_per_batch": round(t_host * 1e3, 2)},
        "plain_per_source": plain,
        "speedup_vs_plain": round(gteps / plain["gteps"], 2) if plain else None,
        "bit_exact_sums_vs_host_driven_and_plain": exact,
        "roofline": {"bound": "hbm", "kernel": "batch replay: k_bp_persistent + the emit launches after it",
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_GBps": round(traffic / (avg_launch_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "traffic_frac": (round(traffic / (avg_launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                      if traffic else None),
                     "algorithmic_bytes_per_launch": st.kernel_bytes[11],
                     "avg_launch_ms": round(avg_launch_ms, 3), "launches": args.steps},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


SERVING_DESC = ("serving: directed web-crawl-shaped R-MAT n=20M (a=.6 b=.15 c=.15, seed 7), queries of 1-256 "
                "sources (log-uniform, seed 8), Poisson arrivals (seed 7)")


def serving_setup(slots: int, words_per_tile: int, tune: bool = True):
    """The serving workload's graph (directed R-MAT, n = 20M, ~300M edges,
    a=.6 b=.15 c=.15, seed 7) resident on the device, a server over it, and
    the candidate source vertices (those with out-edges).  With `tune`, the
    autotuner picks the push/pull switch and pull cut-offs on 512 candidate
    sources (seed 11) before the server starts; the server keeps its own
    tile width (words_per_tile, or its default), which sets how many queued
    rows share a tile."""
    import numpy as np

    import xbmsbfs as xb

    t0 = time.perf_counter()
    g = xb.gen_rmat_capped(25, 20_000_000, 386_000_000, 0.6, 0.15, 0.15, seed=7)
    hg = xb.HGraph(g.row_offsets, g.col_indices, directed=True)
    log(f"serving graph: n={g.n} m={g.m} ({time.perf_counter() - t0:.1f}s)")
    cand = np.flatnonzero(g.degrees() > 0).astype(np.int32)
    tun = {"words_per_tile": words_per_tile} if words_per_tile else {}
    tuned = None
    if tune:
        t1 = time.perf_counter()
        sample = np.random.default_rng(11).choice(cand, size=512, replace=False).astype(np.int32)
        best, rep = xb.autotune(hg, sample, **tun)
        best.pop("words_per_tile", None)
        tuned = {"seconds": round(time.perf_counter() - t1, 2), "trials": rep.trials,
                 "default_ms_per_source": round(rep.default_ms_per_source, 5),
                 "best_ms_per_source": round(rep.best_ms_per_source, 5), "chosen": best}
        log(f"serving autotune: {tuned}")
        tun.update(best)
    srv = xb.Server(hg, slots=slots, **tun)
    srv.tuned = tuned
    return g, hg, srv, cand


def serving_requests(cand, count: int, seed: int):
    """Source sets of 1-256 vertices, log-uniform in size, uniform over cand."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        S = max(1, min(256, int(np.floor(np.exp(rng.uniform(0.0, np.log(257.0)))))))
        out.append(rng.choice(cand, size=S, replace=True).astype(np.int32))
    return out


def serve_once(srv, n: int, reqs, rate: float, seed: int) -> dict:
    """Submits reqs at Poisson arrivals of `rate` per second (seed), waits for
    all, and returns latency and throughput figures.  Latency runs from a
    query's scheduled arrival to the end of its last tile; service time from
    its first tile's start; aggregate GTEPS is the traversed edges of all
    queries over the span from the first arrival to the last completion."""
    import queue
    import threading

    import numpy as np

    N = len(reqs)
    arrivals = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, size=N))
    base = srv.counters()
    infos = [None] * N
    todo = queue.Queue()

    def collect():
        while True:
            item = todo.get()
            if item is None:
                return
            i, rid = item
            inf = srv.wait(rid)
            infos[i] = (inf.t_submit_us, inf.t_admit_us, inf.t_start_us, inf.t_done_us, inf.edges, inf.S, inf.tiles)
            srv.release(rid)

    workers = [threading.Thread(target=collect) for _ in range(16)]
    for w in workers:
        w.start()
    start = time.perf_counter()
    for i in range(N):
        delay = start + arrivals[i] - arrivals[0] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        todo.put((i, srv.submit(reqs[i])))
    for _ in workers:
        todo.put(None)
    for w in workers:
        w.join()
    c = srv.counters()
    t_sub = np.array([x[0] for x in infos])
    sched = t_sub[0] + (arrivals - arrivals[0]) * 1e6  # scheduled arrivals on the server clock
    done = np.array([x[3] for x in infos])
    start_t = np.array([x[2] for x in infos])
    admit = np.array([x[1] for x in infos])
    lat_ms = (done - np.minimum(sched, t_sub)) / 1e3
    svc_ms = (done - start_t) / 1e3
    wait_ms = (admit - t_sub) / 1e3
    edges = int(sum(x[4] for x in infos))
    span_s = (done.max() - sched.min()) / 1e6
    out_bytes = sum(8 * x[5] * n for x in infos)
    tiles = int(c.tiles - base.tiles)
    return {
        "offered_per_s": round(rate, 1),
        "queries": N,
        "sources": int(sum(x[5] for x in infos)),
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 2),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 2),
        "service_p50_ms": round(float(np.percentile(svc_ms, 50)), 2),
        "service_p99_ms": round(float(np.percentile(svc_ms, 99)), 2),
        "admission_wait_p50_ms": round(float(np.percentile(wait_ms, 50)), 2),
        "queries_per_s": round(N / span_s, 1),
        "gteps": round(edges / span_s / 1e9, 1),
        "output_TBps": round(out_bytes / span_s / 1e12, 3),
        "tiles": tiles,
        "rows_per_tile": round((c.tile_rows - base.tile_rows) / max(1, tiles), 1),
        "peak_table_GB": round(c.peak_used_bytes / 1e9, 1),
        "span_ms": round(span_s * 1e3, 1),
    }


def serving_summary(slots: int, words_per_tile: int, requests: int, knee_requests: int, tune: bool = True) -> dict:
    """The serving sub-object of the default bench line: the configured 2,000
    queries/s offered rate (far past one GPU's capacity: latency is admission
    wait), the capacity that run measures, and a sweep of lower offered rates
    for the knee -- the highest rate whose p99 stays within 2x the service
    p99 (the latency of a query that never waits for admission)."""
    g, hg, srv, cand = serving_setup(slots, words_per_tile, tune)
    warm = serving_requests(cand, 3, seed=9)
    serve_once(srv, g.n, warm, 2000.0, seed=9)
    at_2k = serve_once(srv, g.n, serving_requests(cand, requests, seed=8), 2000.0, seed=7)
    cap = at_2k["queries_per_s"]
    log(f"serving at 2000/s: {at_2k}")
    sweep = []
    for k, frac in enumerate((0.5, 0.65, 0.8, 0.95)):
        r = serve_once(srv, g.n, serving_requests(cand, knee_requests, seed=20 + k), frac * cap, seed=30 + k)
        r["within_2x_service_p99"] = r["p99_ms"] <= 2.0 * r["service_p99_ms"]
        sweep.append(r)
        log(f"serving at {frac:.2f} x capacity: {r}")
    ok = [r for r in sweep if r["within_2x_service_p99"]]
    knee = max(ok, key=lambda r: r["offered_per_s"]) if ok else None
    srv.close()
    hg.close()
    return {"workload": SERVING_DESC, "n": g.n, "m": g.m, "slots": slots,
            "words_per_tile": words_per_tile or 8, "autotune": srv.tuned, "at_2000_per_s": at_2k, "capacity_queries_per_s": cap,
            "knee": ({"offered_per_s": knee["offered_per_s"], "p50_ms": knee["p50_ms"], "p99_ms": knee["p99_ms"],
                      "service_p99_ms": knee["service_p99_ms"], "gteps": knee["gteps"]} if knee else None),
            "sweep": sweep}


def bench_serving(args) -> int:
    """Serving on its own line: the web-crawl-shaped graph resident on the
    device, --requests queries at Poisson arrivals of --rate per second."""
    g, hg, srv, cand = serving_setup(args.slots, args.words_per_tile, not args.no_autotune)
    serve_once(srv, g.n, serving_requests(cand, args.warmup, seed=9), args.rate, seed=9)
    r = serve_once(srv, g.n, serving_requests(cand, args.requests, seed=8), args.rate, seed=7)
    srv.close()
    line = {
        "metric": METRIC,
        "value": r["gteps"],
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": r["queries"],
        "warmup": args.warmup,
        "ms_per_step": round(r["span_ms"] / r["queries"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": SERVING_DESC + f", {args.rate:g}/s offered", "n": g.n, "m": g.m,
                   "queries": r["queries"], "slots": args.slots, "words_per_tile": args.words_per_tile or 8,
                   "sources": r["sources"], "autotune": srv.tuned},
        **{k: r[k] for k in ("p50_ms", "p99_ms", "service_p50_ms", "service_p99_ms", "admission_wait_p50_ms",
                             "queries_per_s", "offered_per_s", "output_TBps", "tiles", "rows_per_tile",
                             "peak_table_GB")},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_calibrate(args) -> int:
    """Known-traffic kernels for the PMC byte counters: a linear device copy of
    --copy-calibration-mib MiB (that many bytes read and written) and a gather
    of random 128-byte rows (int32 x 32, line-aligned) from a 4 GiB table into
    a dense output (rows x 128 bytes read and written, plus 8 bytes of index
    per row).  Prints the bytes each kernel must move."""
    import torch

    mib = args.copy_calibration_mib or 1024
    a = torch.ones(mib << 18, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    table = torch.zeros((4 << 30) // 128, 32, dtype=torch.int32, device="cuda")
    rows = (mib << 20) // 128
    idx = torch.randint(0, table.shape[0], (rows,), device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    torch.cuda.synchronize()
    out = torch.index_select(table, 0, idx)
    torch.cuda.synchronize()
    print(json.dumps({"copy": {"read_bytes": mib << 20, "write_bytes": mib << 20},
                      "gather_128B_rows": {"rows": rows, "read_bytes": rows * 128 + rows * 8,
                                           "write_bytes": rows * 128, "table_bytes": table.numel() * 4},
                      "checks": [int(out.shape[0])]}), flush=True)
    return 0


def bench_match(args) -> int:
    """Forest mode: FindMaximumMatch (xb_match_run) from the empty matching on
    the LiveJournal-shaped graph, the shape of the reference's matching test
    on LiveJournal (apps/xb/test/xblsm_gpu/xblsm_gpu_10_test.cu:103-108).  A
    step is one whole matching run, graph resident in HBM, mate array written
    to a device buffer.  GTEPS here = frontier edges read by the scan, bridge
    and claim passes (the reference's num_edges_processed_) over wall time."""
    import numpy as np

    import oracle
    import xbmsbfs as xb

    if args.match_graph == "gnp":
        # Erdos-Renyi G(n, p), n = 1M, mean degree 4 (seed 1): its late
        # phases grow trees over most of the graph and form 235k blossoms.
        n_ = 1_000_000
        rng = np.random.default_rng(1)
        draws = int(rng.binomial(n_ * (n_ - 1) // 2, 4.0 / n_))
        u = rng.integers(0, n_, draws)
        v = rng.integers(0, n_, draws)
        keep = u != v
        a = np.concatenate([u[keep], v[keep]])
        b = np.concatenate([v[keep], u[keep]])
        key = np.unique(a.astype(np.int64) * n_ + b)
        ro = np.zeros(n_ + 1, dtype=np.int64)
        np.add.at(ro, key // n_ + 1, 1)

        class _G:
            pass

        g = _G()
        g.row_offsets = np.cumsum(ro).astype(np.int32)
        g.col_indices = (key % n_).astype(np.int32)
        g.n, g.m = n_, int(key.size // 2)
        desc = "G(n, p) n=1M mean degree 4 (seed 1)"
    else:
        g, _src, _ = make_workload("livejournal")
        desc = "LiveJournal-shaped Chung-Lu (gamma 2.1, seed 4)"
    hg = xb.HGraph(g.row_offsets, g.col_indices)
    out = xb.DeviceBuffer((g.n,))
    regrow = args.match_regrow
    for _ in range(args.warmup):
        xb.match(hg, None, out.ptr, device=True, regrow=regrow)
    t0 = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        _, st = xb.match(hg, None, out.ptr, device=True, regrow=regrow)
        stats.append(st)
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    st = stats[-1]
    mate = out.numpy()
    want_pairs = st.matched_vertices // 2
    assert int(np.count_nonzero(mate >= 0)) == st.matched_vertices
    # Per-class kernel times on one more run (HIP events on the engine's stream).
    _, pst = xb.match(hg, None, out.ptr, device=True, profile=True, regrow=regrow)
    kern = pst.kernels()
    dom = max(kern, key=lambda k: kern[k]["ms"])
    achieved = kern[dom]["bytes"] / (kern[dom]["ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles("match_livejournal", dom), "ms": round(kern[dom]["ms"], 3),
                "launches": kern[dom]["launches"]}
    cpu = None
    bit_exact = None
    if not args.no_cpu:
        # The whole run on one core (about 20 s): the same algorithm, so its
        # matching must equal the GPU's entry for entry.
        t1 = time.perf_counter()
        cmate, cst = oracle.match_forest(g.row_offsets, g.col_indices, reuse=not regrow)
        ct = time.perf_counter() - t1
        bit_exact = bool(np.array_equal(cmate, mate)) and cst.edges_scanned == st.edges_scanned
        cpu = {"value": round(cst.edges_scanned / ct / 1e9, 4), "unit": "GTEPS", "cores": 1, "kind": "port",
               "ms_per_run": round(ct * 1e3, 1),
               "sample": f"the whole matching run (oracle forest restatement, {cst.edges_scanned} edges, "
                         f"{ct:.1f} s)"}
    out.free()
    line = {
        "metric": "FindMaximumMatch (forest mode) GTEPS on 1 MI355X, bit-exact vs the CPU restatement",
        "value": round(st.edges_scanned / (ms_step * 1e-3) / 1e9, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "match: " + desc + ", maximum matching from the empty matching",
                   "n": g.n, "m": g.m, "forest": "regrow every phase" if regrow else "tree reuse"},
        "reused_tree_vertices": int(st.reused_tree_vertices),
        "reset_tree_vertices": int(st.reset_tree_vertices),
        "aug_path_edges": int(st.aug_path_edges),
        "expand_edges": int(st.expand_edges),
        "blossom_scan_edges": int(st.blossom_scan_edges),
        "matched_pairs": int(want_pairs),
        "phases": st.phases,
        "levels": st.levels,
        "augmentations": int(st.augmentations),
        "blossoms": int(st.blossoms),
        "blossom_vertices": int(st.blossom_vertices),
        "bit_exact_vs_cpu": bit_exact,
        "edges_scanned": int(st.edges_scanned),
        "kernels": {k: {"ms": round(v["ms"], 3), "launches": v["launches"],
                        "GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)} for k, v in kern.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="livejournal", choices=["livejournal", "amazon", "rmat16", "uneven", "serving",
                                                                 "match", "calibrate"])
    ap.add_argument("--match-graph", default="livejournal", choices=["livejournal", "gnp"],
                    help="match: the LiveJournal-shaped graph, or a blossom-heavy G(n, p)")
    ap.add_argument("--match-regrow", action="store_true",
                    help="match: regrow the forest every phase instead of keeping the trees of exposed roots")
    ap.add_argument("--requests", type=int, default=400, help="serving: timed queries")
    ap.add_argument("--rate", type=float, default=2000.0, help="serving: Poisson arrival rate (queries/s)")
    ap.add_argument("--slots", type=int, default=2,
                    help="serving: concurrent tiles (2 measured best on the 20M-vertex graph: the table budget caps rows in flight)")
    ap.add_argument("--engine", default="auto", choices=["auto", "hostloop", "persistent"],
                    help="hostloop: one launch per phase, host-driven level loop; persistent: the batch API "
                         "(one persistent kernel, level loop on the device, replayed hipGraph); auto: time "
                         "both with the tuned options and keep the faster")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (words per tile, alpha, pull cut-offs) before the timed region")
    ap.add_argument("--words-per-tile", type=int, default=0)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-rounds", type=int, default=5, help="CPU baseline: rounds per mode and thread count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving sub-object of the default line (web-crawl graph, 2k/s and the knee)")
    ap.add_argument("--serving-requests", type=int, default=200,
                    help="default line: queries of the 2,000/s serving run")
    ap.add_argument("--knee-requests", type=int, default=100,
                    help="default line: queries per offered rate of the knee sweep")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this many MiB (PMC byte-counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the multi-source BFS path shards by source tile with no "
              "per-level exchange, and the engine runs on a single MI355X; use --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import xbmsbfs as xb

    if args.workload == "uneven":
        return bench_uneven(args)
    if args.workload == "match":
        return bench_match(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "calibrate":
        return bench_calibrate(args)

    t0 = time.perf_counter()
    g, src, desc = make_workload(args.workload)
    S, n = int(src.size), g.n
    log(f"workload {args.workload}: n={n} m={g.m} S={S} maxdeg={int(g.degrees().max())} "
        f"({time.perf_counter() - t0:.1f}s)")
    hg = xb.HGraph(g.row_offsets, g.col_indices)
    eng = xb.XBlossomGpuEngine10(hg)
    # The tuner starts from the widest tile the batch fills (one 1,024-source
    # tile on the LiveJournal-shaped workload) and moves off it only for a
    # >= 2% faster trial.
    tun = {"words_per_tile": args.words_per_tile or min(16, 1 << max(0, (S + 63) // 64 - 1).bit_length())}
    tune_info = None
    if not args.no_autotune:
        t1 = time.perf_counter()
        best, rep = xb.autotune(hg, src, **tun)
        tune_info = {"seconds": round(time.perf_counter() - t1, 2), "trials": rep.trials, "passes": rep.passes,
                     "default_ms_per_source": round(rep.default_ms_per_source, 5),
                     "best_ms_per_source": round(rep.best_ms_per_source, 5), "chosen": best}
        tun = best
        log(f"autotune: {tune_info}")
    lv = torch.empty((S, n), dtype=torch.int32, device="cuda")
    pa = torch.empty((S, n), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    batch = None
    engine = args.engine
    if engine in ("persistent", "auto"):
        batch = xb.Batch([hg], [src], [lv.data_ptr()], [pa.data_ptr()], profile=True, **tun)
    engine_ms = {}
    if engine == "auto":
        for name in ("hostloop", "persistent"):
            best_ms = 1e30
            for _ in range(3):
                if name == "persistent":
                    best_ms = min(best_ms, batch.run().ms_total)
                else:
                    eng.MultiSourceBFS(src, lv.data_ptr(), pa.data_ptr(), mode=xb.XB_MODE_BATCHED, device_out=True,
                                       **tun)
                    best_ms = min(best_ms, eng.last_stats.ms_total)
            engine_ms[name] = round(best_ms, 3)
        engine = min(engine_ms, key=engine_ms.get)
        log(f"engine choice: {engine_ms} -> {engine}")
    if engine != "persistent" and batch is not None:
        batch.close()
        batch = None

    def run_batched(profile: bool):
        if batch is not None:
            return batch.run()
        eng.MultiSourceBFS(src, lv.data_ptr(), pa.data_ptr(), mode=xb.XB_MODE_BATCHED, device_out=True,
                           profile=profile, **tun)
        return eng.last_stats

    for _ in range(args.warmup):
        run_batched(False)
    torch.cuda.synchronize()
    log("warmup done")

    # Timed region: K batched steps, per-kernel HIP events on the engine's stream.
    agg_ms = {}
    agg_bytes = {}
    agg_launch = {}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    last = None
    for _ in range(args.steps):
        st = run_batched(True)
        for name, k in st.kernels().items():
            agg_ms[name] = agg_ms.get(name, 0.0) + k["ms"]
            agg_bytes[name] = agg_bytes.get(name, 0.0) + k["bytes"]
            agg_launch[name] = agg_launch.get(name, 0) + k["launches"]
        last = st.as_dict()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t_start
    # Several tiles per step run on two streams (the tile pipeline), so a
    # kernel's HIP-event span would include the other stream's work: the
    # per-kernel figures then come from as many profiled steps on one stream.
    wpt_run = int(tun.get("words_per_tile", 8))
    kernels_source = "timed steps"
    if batch is None and S > 64 * wpt_run:
        agg_ms, agg_bytes, agg_launch = {}, {}, {}
        for _ in range(args.steps):
            eng.MultiSourceBFS(src, lv.data_ptr(), pa.data_ptr(), mode=xb.XB_MODE_BATCHED, device_out=True,
                               profile=True, pipeline=False, **tun)
            for name, k in eng.last_stats.kernels().items():
                agg_ms[name] = agg_ms.get(name, 0.0) + k["ms"]
                agg_bytes[name] = agg_bytes.get(name, 0.0) + k["bytes"]
                agg_launch[name] = agg_launch.get(name, 0) + k["launches"]
        torch.cuda.synchronize()
        kernels_source = "one-stream profiled steps after the timed region (the timed steps pipeline tiles on two streams)"
    edges = eng.CountTraversed(lv.data_ptr(), S)
    gteps = edges * args.steps / t_batched / 1e9
    ms_per_step = t_batched / args.steps * 1e3
    log(f"batched: {ms_per_step:.2f} ms/step, {gteps:.1f} GTEPS, edges/step={edges}")

    if args.copy_calibration_mib:
        src_t = torch.ones(args.copy_calibration_mib * (1 << 18), dtype=torch.int32, device="cuda")
        dst_t = torch.empty_like(src_t)
        torch.cuda.synchronize()
        dst_t.copy_(src_t)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_calibration_mib} MiB read + {args.copy_calibration_mib} MiB written")
        del src_t, dst_t

    # Plain per-source engine over the same batch.
    plain = None
    if not args.no_plain:
        lv2 = torch.empty((S, n), dtype=torch.int32, device="cuda")
        pa2 = torch.empty((S, n), dtype=torch.int32, device="cuda")
        eng.MultiSourceBFS(src[:4], lv2.data_ptr(), pa2.data_ptr(), mode=xb.XB_MODE_PLAIN, device_out=True)
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        for _ in range(args.plain_steps):
            eng.MultiSourceBFS(src, lv2.data_ptr(), pa2.data_ptr(), mode=xb.XB_MODE_PLAIN, device_out=True,
                               profile=True)
        torch.cuda.synchronize()
        t_plain = (time.perf_counter() - tp0) / args.plain_steps
        exact = bool(torch.equal(lv, lv2)) and bool(torch.equal(pa, pa2))
        plain = {"gteps": round(edges / t_plain / 1e9, 3), "ms_per_batch": round(t_plain * 1e3, 2),
                 "bit_exact_vs_batched": exact, "kernels": eng.last_stats.kernels()}
        log(f"plain: {t_plain * 1e3:.1f} ms/batch, {plain['gteps']} GTEPS, bit-exact={exact}")
        del lv2, pa2

    # Roofline of the dominant kernel class (largest summed HIP-event time).
    dom = max(agg_ms, key=lambda k: agg_ms[k])
    avg_ms = agg_ms[dom] / agg_launch[dom]
    bytes_per_launch = agg_bytes[dom] / agg_launch[dom]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    gsum = f"{g.checksum():016x}"
    prof_vals = profile_values(args.workload, dom, gsum)
    traffic = float(prof_vals["traffic_bytes_per_launch"]) if "traffic_bytes_per_launch" in prof_vals else None
    roofline = {
        "bound": "hbm",
        "kernel": KERNEL_SYMBOL.get(dom, dom),
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        # The same launch's measured bytes (committed PMC summary of these
        # engine sources and this graph) over its measured time: an upper
        # bound on its HBM traffic, beside the algorithmic figure above.
        "traffic_GBps": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
        "traffic_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
        "traffic_source": TRAFFIC_SOURCE if traffic else prof_vals.get("_stale"),
        "algorithmic_bytes_per_launch": bytes_per_launch,
        "avg_launch_ms": round(avg_ms, 4),
        "launches": agg_launch[dom],
        "measured_in": kernels_source,
    }
    # Whole-step view: the step's algorithmic bytes (CSR and state streams per
    # level, plus the output tables) over the step time.
    step_bytes = sum(agg_bytes.values()) / args.steps
    # Whole-step measured traffic, when this run is in the profiled
    # configuration (engine sources, graph, engine and words per tile): every
    # engine kernel and fill of one step, from the committed PMC summary.
    step_traffic = None
    wpt = tun.get("words_per_tile", args.words_per_tile or 8)
    if ("step_traffic_bytes" in prof_vals and prof_vals.get("profiled_engine") == engine
            and prof_vals.get("profiled_words_per_tile") == str(wpt)):
        tb = float(prof_vals["step_traffic_bytes"])
        step_traffic = {"bytes": tb, "GBps": round(tb / (ms_per_step * 1e-3) / 1e9, 1),
                        "frac": round(tb / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "source": TRAFFIC_SOURCE}
    # Gather-aware bound of the pull: every pull level may read each in-edge's
    # source frontier line (8 B per word) besides the CSR and the per-vertex
    # state the algorithmic count holds; the early exit reads fewer.
    T_used = int(wpt)
    pull_levels = last["pull_levels"] if last else 0
    pull_gather_bound = pull_levels * ((4.0 + 8.0 * T_used) * g.m + 4.0 * (n + 1) + 24.0 * T_used * n)
    kernels = {k: {"ms_per_step": round(agg_ms[k] / args.steps, 3),
                   "launches_per_step": agg_launch[k] / args.steps,
                   "GBps": round(agg_bytes[k] / (agg_ms[k] * 1e-3) / 1e9, 1) if agg_ms[k] > 0 else None}
               for k in sorted(agg_ms, key=lambda k: -agg_ms[k])}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, src, args.cpu_rounds)
        log(f"cpu baseline: {cpu}")

    # Serving on the web-crawl-shaped graph: the LiveJournal-shaped tables are freed first.
    serving = None
    if not args.no_serving and args.workload == "livejournal":
        if batch is not None:
            batch.close()
        del lv, pa
        torch.cuda.empty_cache()
        hg.close()
        serving = serving_summary(args.slots, 0, args.serving_requests, args.knee_requests, not args.no_autotune)

    line = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "n": n,
            "m": int(g.m),
            "sources": S,
            "graph_checksum": f"{g.checksum():016x}",
            "edges_traversed_per_step": int(edges),
            "max_level": last["max_level"] if last else None,
            "words_per_tile": tun.get("words_per_tile", args.words_per_tile or 8),
            "engine": engine,
            "engine_ms": engine_ms or None,
            "autotune": tune_info,
        },
        "plain_per_source": plain,
        "speedup_vs_plain": round(gteps / plain["gteps"], 2) if plain else None,
        "roofline": roofline,
        "step_algorithmic_GBps": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "step_algorithmic_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "pull_gather_bound_bytes_per_step": pull_gather_bound,
        "step_traffic": step_traffic,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
