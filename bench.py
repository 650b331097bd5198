"""Headline benchmark: updated-edge GTEPS, batched vs plain, on one MI355X.

Default workload (the metric's configuration, C3): PageRank and CC labels on a
seeded R-MAT graph of 2^27 vertices, edge factor 32 (a/b/c/d =
0.57/0.19/0.19/0.05, self-loops and duplicates dropped: 4.22 billion edges),
the edge arrays in pinned host memory under an 8 GiB device edge budget, the
-cache 4 HBM edge cache, SEGMENT 512; every snapshot applies 1,000,000 fresh
inserts (80 % from the top 1 % out-degree sources) + 1,000,000 deletes
sampled from the current edges.

  plain    one engine per program (PageRank, CC), each with its own stores
  batched  one engine for both programs on one forward and one in-edge store
           (SG_ALGO_PR_CC): each batch updates each store once, CC's recompute
           reads the in-lists of its reset vertices

  updated-edge GTEPS = sum_t (E_t x programs answered on G_t) / sum_t T_t / 1e9
  T_t = wall time from hand-off of the host batch (sg_apply_snapshot: upload,
        validation, PMA updates, recompute and cache maintenance) to converged
        device state

A C3 snapshot costs tens of seconds (PageRank's ~100 rounds at epsilon 1e-6
each read the 4.2-billion-edge in-store, mostly over the host link), so the
C3 part times --c3-steps snapshots per path (default 2, after --c3-warmup);
--steps/--warmup apply to the secondary C2 line (incremental SSSP + BFS from
source 0 on R-MAT scale 22 EF16, 100k + 100k per snapshot, HBM-resident),
printed in the same JSON object under "c2", and a C4 key: 64 BFS + 64 SSSP
queries on R-MAT scale 24 EF16 under batches of log-uniform size in [1e3, 1e6],
the batched engine against single-source engines on a stratified sample of
the same sources (--no-c4 skips it).

Prints one JSON line. `--gpus` above 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "c-gpustreamgraph_amd"))

import numpy as np  # noqa: E402

import sgraph  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel, summary="k_sweep_edges_sssp_bfs_rmat22_ef16_b100k.pmc.txt"):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 PMC summary
    under profiles/ (FETCH_SIZE + WRITE_SIZE, calibrated against a plain
    copy), or None."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None, None
    vals = []
    for line in open(path):
        if line.startswith(kernel):
            vals.append(float(line.split()[-1]) * 1e6)
    return (round(sum(vals) / len(vals)), os.path.relpath(path, ROOT)) if vals else (None, None)


def cpu_baseline(row, col, seed, budget_s):
    """The reference's CPU validation path (SSSPHostNaive + BFSHost, restated in
    oracle/oracle.c) from scratch on the snapshot graph, one core: repeated
    until about budget_s seconds of CPU work (at least 3 times), median per
    solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.lib()
    ne = len(col)
    ts, tb = [], []
    t_start = time.perf_counter()
    while len(ts) < 3 or (time.perf_counter() - t_start < budget_s and len(ts) < 20):
        t0 = time.perf_counter()
        O.sssp_fifo(row, col, 0, "seeded", seed)
        t1 = time.perf_counter()
        O.bfs(row, col, 0)
        t2 = time.perf_counter()
        ts.append(t1 - t0)
        tb.append(t2 - t1)
    s_med, b_med = float(np.median(ts)), float(np.median(tb))
    return {
        "value": round(2 * ne / (s_med + b_med) / 1e9, 6),
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{len(ts)} x (one from-scratch SSSP (FIFO Bellman-Ford, SSSPHostNaive) + one BFS (BFSHost)) on G_0 "
                  f"({len(row) - 1} vertices, {ne} edges), {time.perf_counter() - t_start:.1f} s in all; "
                  f"medians {s_med:.2f} s + {b_med:.2f} s",
        "host_cpus": os.cpu_count(),
    }


def run_batched(args, row, col, ins, dele, wseed, algos, plain_value):
    """The same queries answered together on one shared graph (one PMA update,
    one boundary sweep and one launch per round for all of them)."""
    nsnap = args.warmup + args.steps
    e = sgraph.Engine("multi", weight_mode="seeded", weight_seed=wseed, rounds_per_sync=args.rounds_per_sync,
                      flags=sgraph.FLAG_GRAPH_ROUNDS if args.graph_rounds else 0)
    t = time.perf_counter()
    e.load_csr(row, col)
    e.init_queries([0] * len(algos), algos)
    log(f"[bench] batched {'+'.join(algos)}: load+static solve {time.perf_counter() - t:.2f}s")
    for s in range(args.warmup):
        e.apply(ins[s], dele[s])
    e.synchronize()
    lat, num = [], 0
    for s in range(args.warmup, nsnap):
        t0 = time.perf_counter()
        e.apply(ins[s], dele[s])
        lat.append(time.perf_counter() - t0)
        num += e.stats()["ne"] * len(algos)
    value = num / sum(lat) / 1e9
    e.close()
    lat_ms = np.array(lat) * 1e3
    return {
        "value": round(value, 4), "unit": "GTEPS", "vs_plain": round(value / plain_value, 3),
        "latency_p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "latency_p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "what": f"{len(algos)} queries ({'+'.join(algos)} from source 0) on one shared graph: one PMA update, one "
                f"sweep and one launch per round per snapshot (sg_init_queries)",
    }


def run_oom(args):
    """Out-of-memory path (PageRank + CC with the edge array in pinned host
    memory under an edge budget, the HBM edge cache and SEGMENT streaming):
    updated-edge GTEPS, host-link GB/s of the streamed segments, cache hit rate."""
    nsnap = args.warmup + args.steps
    t = time.perf_counter()
    nv, row, col = sgraph.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    ins, dele = sgraph.gen_updates(row, col, args.scale, nsnap, args.batch, args.batch, seed=args.seed + 1,
                                   hub_frac=0.8, hub_top=0.01)
    log(f"[bench] oom graph nv={nv} ne={len(col)} gen {time.perf_counter() - t:.1f}s")
    res = {}
    for a in args.algos.split(","):
        kw = {"pma_density": args.pma_density} if args.pma_density else {}
        e = sgraph.Engine(a, edge_budget_gb=args.edge_budget_gb, cache_level=args.cache, segments=args.segments, **kw)
        t = time.perf_counter()
        e.load_csr(row, col)
        e.init()
        st = e.stats()
        log(f"[bench] oom {a}: host_edges={st['host_edges']} load+static solve {time.perf_counter() - t:.1f}s "
            f"cache={st['cache_slots']} cached_vertices={st['cached_vertices']}")
        for s in range(args.warmup):
            e.apply(ins[s], dele[s])
        e.reset_stats()
        lat, num = [], 0
        for s in range(args.warmup, nsnap):
            t0 = time.perf_counter()
            e.apply(ins[s], dele[s])
            lat.append(time.perf_counter() - t0)
            num += e.stats()["ne"]
            log(f"[bench] oom {a} snapshot {s}: {lat[-1] * 1e3:.1f} ms")
        st = e.stats()
        e.close()
        res[a] = {"gteps": num / sum(lat) / 1e9, "lat_ms": [x * 1e3 for x in lat], "stats": st}
    num = sum(sum(x["lat_ms"]) for x in res.values())
    edges = 0
    for a, x in res.items():
        edges += x["gteps"] * sum(x["lat_ms"]) * 1e6
    value = edges / (num * 1e-3) / 1e9
    sb = sum(x["stats"]["streamed_bytes"] for x in res.values())
    sms = sum(x["stats"]["stream_ms"] for x in res.values())
    cfb = sum(x["stats"]["cache_fill_bytes"] for x in res.values())
    cfms = sum(x["stats"]["cache_fill_ms"] for x in res.values())
    rel = sum(x["stats"]["edges_relaxed"] for x in res.values())
    hit = sum(x["stats"]["edges_from_cache"] for x in res.values())
    all_lat = np.array([y for x in res.values() for y in x["lat_ms"]])
    return {
        "metric": "updated-edge GTEPS (incremental PR+CC, out-of-memory path: edges in pinned host memory)",
        "value": round(value, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "dtype": "u32/f32", "data": "synthetic (seeded R-MAT, hub-biased update stream)",
        "config": {"workload": f"PR+CC on R-MAT scale {args.scale} EF{args.edge_factor}, edge budget "
                               f"{args.edge_budget_gb} GiB, -cache {args.cache}, SEGMENT {args.segments}, "
                               f"{args.batch}+{args.batch} updates/snapshot (80% of inserts from the top-1% sources)"
                               + (f", PMA density {args.pma_density}" if args.pma_density else ""),
                   "nv": nv, "ne0": len(col)},
        "latency_p50_ms": round(float(np.percentile(all_lat, 50)), 3),
        "latency_p99_ms": round(float(np.percentile(all_lat, 99)), 3),
        "host_link": {"streamed_bytes": sb, "stream_ms": round(sms, 3),
                      "achieved_gbs": round(sb / (sms * 1e-3) / 1e9, 2) if sms else None,
                      "achieved_what": "segment bytes of the streamed sweeps / their wall time (copies and sweeps)",
                      "cache_fill_bytes": cfb, "cache_fill_ms": round(cfms, 3),
                      "cache_fill_gbs": round(cfb / (cfms * 1e-3) / 1e9, 2) if cfms else None, "peak_gbs": 64.0},
        "cache_hit_rate": round(hit / rel, 4) if rel else None,
        "per_algo": {a: {"gteps": round(x["gteps"], 4), "p50_ms": round(float(np.percentile(x["lat_ms"], 50)), 3),
                         "host_edges": x["stats"]["host_edges"], "cached_vertices": x["stats"]["cached_vertices"],
                         "host_bytes": x["stats"]["host_bytes"], "device_bytes": x["stats"]["device_bytes"]}
                     for a, x in res.items()},
    }


def run_batch(args):
    """Many BFS + SSSP sources on one graph under a stream of update batches of
    log-uniform sizes: the batched engine (all queries, one PMA update, one
    sweep and one launch per round) against plain single-source engines."""
    rng = np.random.default_rng(args.seed)
    t = time.perf_counter()
    nv, row, col = sgraph.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    nsnap = args.warmup + args.steps
    ins, dele = sgraph.gen_updates(row, col, args.scale, nsnap, args.max_update // 2, args.max_update // 2,
                                   seed=args.seed + 1)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(args.max_update), nsnap)).astype(np.int64)
    batches = [(ins[i][: sizes[i] // 2], dele[i][: sizes[i] - sizes[i] // 2]) for i in range(nsnap)]
    nq = args.queries
    sources = rng.integers(0, nv, nq).astype(np.uint32)
    kinds = ["bfs"] * (nq // 2) + ["sssp"] * (nq - nq // 2)
    wseed = args.seed + 2
    log(f"[bench] batch graph nv={nv} ne={len(col)} queries={nq} sizes={sizes.tolist()} gen {time.perf_counter() - t:.1f}s")

    def run(engine_algo, init, reached=False):
        e = sgraph.Engine(engine_algo, weight_mode="seeded", weight_seed=wseed)
        e.load_csr(row, col)
        init(e)
        for s in range(args.warmup):
            e.apply(*batches[s])
        e.synchronize()
        lat, num = [], 0
        for s in range(args.warmup, nsnap):
            t0 = time.perf_counter()
            e.apply(*batches[s])
            lat.append(time.perf_counter() - t0)
            num += e.stats()["ne"]
        if reached:  # vertices the single-source query reaches on the last snapshot
            r = int((e.values() != sgraph.INF).sum())
            e.close()
            return num, lat, r
        if engine_algo == "multi":  # reach of the sampled queries in the batch
            reach_b.extend(int((e.query(q) != sgraph.INF).sum()) for q in sample)
        e.close()
        return num, lat

    reach_b = []
    k = max(2, min(args.plain_sample, nq))
    sample = sorted(set(int(i) for i in np.linspace(0, nq - 1, k).round()))

    num, lat_b = run("multi", lambda e: e.init_queries(sources, kinds))
    batched = num * nq / sum(lat_b) / 1e9
    # plain: single-source engines on the same batches, from a fixed stratified
    # sample of the batch's sources (every (nq / k)-th query, both kinds), so the
    # two sides answer queries of the same reach distribution
    plain_t, plain_e, reach_p = 0.0, 0, []
    for i in sample:
        n1, lat1, reach = run(kinds[i], lambda e, s=int(sources[i]): e.init(s), reached=True)
        plain_t += sum(lat1)
        plain_e += n1
        reach_p.append(reach)
    plain = plain_e / plain_t / 1e9
    lat_ms = np.array(lat_b) * 1e3
    return {
        "metric": "updated-edge GTEPS of batched BFS+SSSP queries (E_t x queries / T_t), 1 MI355X",
        "value": round(batched, 4), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "dtype": "u32", "data": "synthetic (seeded R-MAT, log-uniform batch sizes)",
        "config": {"workload": f"{nq // 2} BFS + {nq - nq // 2} SSSP queries (seeded uniform sources) on R-MAT scale "
                               f"{args.scale} EF{args.edge_factor}, batches log-uniform 1000..{args.max_update} pairs "
                               f"(50/50 insert/delete)", "nv": nv, "ne0": len(col)},
        "latency_p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "latency_p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "plain": {"value": round(plain, 4),
                  "what": f"single-source engines on the same batches, from {len(sample)} of the {nq} sources "
                          f"(evenly spaced over the query list, both kinds)",
                  "reached_mean": round(float(np.mean(reach_p)), 1)},
        "batched_reached_mean": round(float(np.mean(reach_b)), 1) if reach_b else None,  # same sampled queries
        "vs_plain": round(batched / plain, 2),
    }


def run_serve(args):
    """Request loop: Poisson update batches and mixed queries against one
    resident graph; throughput and p50/p99 request latency."""
    import threading
    rng = np.random.default_rng(args.seed)
    t = time.perf_counter()
    nv, row, col = sgraph.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    pool = 32
    ins, dele = sgraph.gen_updates(row, col, args.scale, pool, args.max_update // 2, args.max_update // 2,
                                   seed=args.seed + 1)
    log(f"[bench] serve graph nv={nv} ne={len(col)} gen {time.perf_counter() - t:.1f}s")
    srv = sgraph.Server(row, col, query_slots=args.slots, weight_mode="seeded", weight_seed=args.seed + 2)
    log(f"[bench] serve ready {time.perf_counter() - t:.1f}s")
    dur = args.duration
    lock = threading.Lock()
    done = {"updates": 0, "queries": 0, "lookups": 0}

    def updates():
        r = np.random.default_rng(args.seed + 3)
        t_end = time.perf_counter() + dur
        k = 0
        while time.perf_counter() < t_end:
            time.sleep(r.exponential(1.0 / args.update_rate))
            size = int(np.exp(r.uniform(np.log(100), np.log(args.max_update))))
            b = k % pool
            srv.wait(srv.submit(sgraph.REQ_UPDATE, ins=ins[b][: size // 2], dele=dele[b][: size - size // 2]))
            k += 1
            with lock:
                done["updates"] += 1

    def queries(cid):
        r = np.random.default_rng(args.seed + 10 + cid)
        t_end = time.perf_counter() + dur
        pend = []
        while time.perf_counter() < t_end:
            time.sleep(r.exponential(args.clients / args.query_rate))
            x = r.random()
            kind = sgraph.REQ_BFS if x < 0.4 else sgraph.REQ_SSSP if x < 0.7 else sgraph.REQ_CC if x < 0.9 else sgraph.REQ_PR_TOPK
            arg = int(r.integers(0, nv)) if kind != sgraph.REQ_PR_TOPK else 10
            tgt = int(r.integers(0, nv)) if kind in (sgraph.REQ_BFS, sgraph.REQ_SSSP) else sgraph.INF
            pend.append((kind, arg, tgt, srv.submit(kind, arg, target=tgt)))
            while len(pend) > 64:
                kk, aa, tt, tk = pend.pop(0)
                srv.wait(tk, kk, aa, tt)
        for kk, aa, tt, tk in pend:
            srv.wait(tk, kk, aa, tt)

    th = [threading.Thread(target=updates)] + [threading.Thread(target=queries, args=(i,)) for i in range(args.clients)]
    t0 = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return {
        "metric": "request loop: answered-query GTEPS and p50/p99 request latency (1 MI355X)",
        "value": round(st["edges_answered"] / wall / 1e9, 4), "unit": "GTEPS", "n_gpus": 1,
        "higher_is_better": True, "dtype": "u32/f32", "data": "synthetic (seeded R-MAT, Poisson arrivals)",
        "config": {"workload": f"serving on R-MAT scale {args.scale} EF{args.edge_factor}: updates {args.update_rate}/s "
                               f"(log-uniform 100..{args.max_update} pairs), queries {args.query_rate}/s "
                               f"(40% BFS, 30% SSSP, 20% CC label, 10% PR top-10), {args.slots} query slots, "
                               f"{dur:.0f} s", "nv": nv, "ne0": len(col)},
        "requests_per_s": round((st["queries"] + st["lookups"] + st["snapshots"]) / wall, 1),
        "latency_p50_ms": round(st["p50_ms"], 3), "latency_p99_ms": round(st["p99_ms"], 3),
        "query_p50_ms": round(st["query_p50_ms"], 3), "query_p99_ms": round(st["query_p99_ms"], 3),
        "update_p50_ms": round(st["update_p50_ms"], 3), "update_p99_ms": round(st["update_p99_ms"], 3),
        "server": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()},
    }


def run_c2(args):
    """Incremental SSSP + BFS from source 0 (C2): the plain per-snapshot path on
    the resident configuration, with --steps / --warmup snapshots, and the same
    queries batched on one shared graph."""
    algos = args.algos.split(",")
    nsnap = args.warmup + args.steps
    t = time.perf_counter()
    nv, row, col = sgraph.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    ins, dele = sgraph.gen_updates(row, col, args.scale, nsnap, args.batch, args.batch, seed=args.seed + 1)
    log(f"[bench] graph nv={nv} ne={len(col)} snapshots={nsnap} batch={args.batch} gen {time.perf_counter() - t:.1f}s")
    wseed = args.seed + 2


    engines = {}
    for a in algos:
        timing = 0 if args.no_timing else sgraph.FLAG_KERNEL_TIMING | (0 if args.all_kernel_timing else sgraph.FLAG_TIME_SWEEP)
        flags = timing | (sgraph.FLAG_GRAPH_ROUNDS if args.graph_rounds else 0) | (sgraph.FLAG_IN_EDGES if args.in_edges else 0)
        e = sgraph.Engine(a, weight_mode="seeded", weight_seed=wseed, flags=flags, wave_budget=args.wave_budget,
                          heavy_degree=args.heavy_degree, round_blocks=args.round_blocks,
                          rounds_per_sync=args.rounds_per_sync)
        t = time.perf_counter()
        e.load_csr(row, col)
        e.init(0)
        st = e.stats()
        log(f"[bench] {a}: load+static solve {time.perf_counter() - t:.2f}s rounds={st['rounds']}")
        engines[a] = e

    ne_t = len(col)
    edges_per_step = []
    lat = {a: [] for a in algos}
    for s in range(args.warmup):
        for a in algos:
            engines[a].apply(ins[s], dele[s])
    for a in algos:
        engines[a].reset_stats()
        engines[a].synchronize()
    t_start = time.perf_counter()
    for s in range(args.warmup, nsnap):
        e_t = 0
        for a in algos:
            e = engines[a]
            t0 = time.perf_counter()
            e.apply(ins[s], dele[s])  # host hand-off (upload included); returns at converged device state
            lat[a].append(time.perf_counter() - t0)
            e_t += e.stats()["ne"]
        edges_per_step.append(e_t)
    for a in algos:
        engines[a].synchronize()
    t_total = time.perf_counter() - t_start
    t_sum = sum(sum(v) for v in lat.values())

    value = sum(edges_per_step) / t_sum / 1e9
    all_lat = np.array([x for a in algos for x in lat[a]]) * 1e3
    stats = {a: engines[a].stats() for a in algos}
    # HIP-event time and algorithmic bytes per kernel class, both engines
    kt = {}
    for a in algos:
        for cls, (n, ms, b) in engines[a].kernel_times().items():
            x = kt.setdefault(cls, [0, 0.0, 0])
            x[0] += n
            x[1] += ms
            x[2] += b
    kernels = {cls: {"launches": n, "ms_per_step": round(ms / args.steps, 4),
                     "avg_launch_us": round(ms * 1e3 / n, 2) if n else None,
                     "achieved_gbs": round(b / (ms * 1e-3) / 1e9, 2) if ms and b else None}
               for cls, (n, ms, b) in kt.items()}
    # the roofline is quoted for the class with the most kernel time that has a byte model
    timed = [c for c in kt if kt[c][2] > 0 and kt[c][1] > 0]
    dom = max(timed, key=lambda c: kt[c][1]) if timed else None  # None under --no-timing
    dn, dms, db = kt[dom] if dom else (0, 0.0, 0)
    achieved = db / (dms * 1e-3) / 1e9 if dms > 0 else None
    traffic, traffic_src = (args.traffic, "--traffic") if args.traffic else (
        measured_traffic("sg::k_sweep_edges") if dom == "sweep" and args.scale == 22 and args.edge_factor == 16
        else (None, None))
    names = {"sweep": "k_sweep_edges (boundary sweep over all live edges: 4 B per live edge + 16 B per vertex)",
             "relax": "k_round (convergence rounds: 12 B per relaxed edge + 32 B per expanded vertex)"}
    out = {
        "metric": "updated-edge GTEPS (incremental SSSP+BFS per snapshot, 1 MI355X, plain per-snapshot path; "
                  "T_t from host hand-off)",
        "value": round(value, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_total * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded R-MAT graph and update stream generated in-process)",
        "config": {
            "workload": f"incremental {'+'.join(x.upper() for x in algos)} from source 0 on R-MAT scale {args.scale} "
                        f"EF{args.edge_factor}, {args.batch} inserts + {args.batch} deletes per snapshot, HBM-resident",
            "scale": args.scale, "edge_factor": args.edge_factor, "nv": nv, "ne0": ne_t,
            "batch_inserts": args.batch, "batch_deletes": args.batch, "weights": "seeded uniform [1,63]",
            "source": 0, "queries_per_snapshot": len(algos),
        },
        "snapshot_latency_ms": {a: {"p50": round(float(np.percentile(np.array(lat[a]) * 1e3, 50)), 3),
                                    "p99": round(float(np.percentile(np.array(lat[a]) * 1e3, 99)), 3)} for a in algos},
        "latency_p50_ms": round(float(np.percentile(all_lat, 50)), 3),
        "latency_p99_ms": round(float(np.percentile(all_lat, 99)), 3),
        "roofline": {
            "kernel": names.get(dom, dom),
            "bound": "hbm",
            "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "launches": dn,
            "avg_launch_us": round(dms * 1e3 / dn, 2) if dn else None,
            "algorithmic_bytes_per_launch": round(db / dn) if dn else None,
        },
        "kernels": kernels,
        "engine": {a: {k: stats[a][k] for k in ("rounds", "invalidated", "rebalances", "resizes", "edges_relaxed",
                                               "vertices_processed", "device_bytes")} for a in algos},
    }
    if not args.no_batched:
        out["batched"] = run_batched(args, row, col, ins, dele, wseed, algos, value)
    out["_graph"] = (row, col, wseed)
    for a in algos:
        engines[a].close()
    return out


C3 = {"scale": 27, "edge_factor": 32, "batch": 1000000, "edge_budget_gb": 8.0, "cache": 4, "segments": 512,
      "pma_density": 0.6, "hub_frac": 0.8, "hub_top": 0.01}


def run_c3_engine(algo, row, col, ins, dele, warm, steps, cfg):
    """One engine on the C3 stream: load, static solve, `warm` untimed then
    `steps` timed snapshots from host batches (T_t includes the upload)."""
    e = sgraph.Engine(algo, flags=sgraph.FLAG_KERNEL_TIMING, **cfg)
    t = time.perf_counter()
    e.load_csr(row, col)
    t_load = time.perf_counter() - t
    t = time.perf_counter()
    e.init()
    t_init = time.perf_counter() - t
    st = e.stats()
    log(f"[bench] c3 {algo}: load {t_load:.1f}s static solve {t_init:.1f}s rounds={st['rounds']} "
        f"host_bytes={st['host_bytes'] / 1e9:.1f}GB device_bytes={st['device_bytes'] / 1e9:.1f}GB")
    for s in range(warm):
        e.apply(ins[s], dele[s])
    e.reset_stats()
    lat, edges = [], []
    for s in range(warm, warm + steps):
        t0 = time.perf_counter()
        e.apply(ins[s], dele[s])  # returns at converged device state
        lat.append(time.perf_counter() - t0)
        st = e.stats()
        edges.append(st["ne"])
        log(f"[bench] c3 {algo} snapshot {s}: {lat[-1]:.2f}s rounds={st['rounds']} pulls={st['pull_rounds']} "
            f"invalidated={st['invalidated']}")
    st = e.stats()
    kt = e.kernel_times()
    e.close()
    return {"lat": lat, "edges": edges, "stats": st, "kt": kt, "load_s": t_load, "init_s": t_init}


def pct(x, p):
    return round(float(np.percentile(np.array(x) * 1e3, p)), 1)


def run_c3(args):
    """C3: PageRank + CC, plain (one engine each) against batched (SG_ALGO_PR_CC)."""
    warm, steps = args.c3_warmup, args.c3_steps
    t = time.perf_counter()
    nv, row, col = sgraph.gen_rmat_device(C3["scale"], C3["edge_factor"], seed=args.seed)
    t_gen = time.perf_counter() - t
    t = time.perf_counter()
    ins, dele = sgraph.gen_updates(row, col, C3["scale"], warm + steps, C3["batch"], C3["batch"], seed=args.seed + 1,
                                   hub_frac=C3["hub_frac"], hub_top=C3["hub_top"])
    log(f"[bench] c3 graph nv={nv} ne={len(col)} gen {t_gen:.1f}s updates {time.perf_counter() - t:.1f}s")
    cfg = {"edge_budget_gb": C3["edge_budget_gb"], "cache_level": C3["cache"], "segments": C3["segments"],
           "pma_density": C3["pma_density"]}
    res = {a: run_c3_engine(a, row, col, ins, dele, warm, steps, cfg) for a in ("cc", "pr", "prcc")}
    # plain: each program on its own engine; batched: both on one engine
    t_plain = sum(res["pr"]["lat"]) + sum(res["cc"]["lat"])
    e_plain = sum(res["pr"]["edges"]) + sum(res["cc"]["edges"])
    t_b = sum(res["prcc"]["lat"])
    e_b = 2 * sum(res["prcc"]["edges"])
    plain, batched = e_plain / t_plain / 1e9, e_b / t_b / 1e9
    sb = res["prcc"]["stats"]
    n_pull, ms_pull, b_pull = res["prcc"]["kt"]["pull"]
    pulls = sb["pull_rounds"]
    # bytes of a pull that cross the host link: the live in-edges outside the HBM
    # copy, read from the packed host copy (ceil(log2 nv) bits per id)
    id_bits = max(1, (nv - 1).bit_length())
    host_per_pull = (sb["ne"] - sb["in_mirror_live"]) * id_bits / 8
    achieved = b_pull / (ms_pull * 1e-3) / 1e9 if ms_pull else None
    return {
        "nv": nv, "ne0": len(col), "plain": plain, "batched": batched, "res": res, "pulls": pulls,
        "ms_pull": ms_pull, "b_pull": b_pull, "achieved": achieved, "host_per_pull": host_per_pull,
        "lat_b": res["prcc"]["lat"], "steps": steps, "warm": warm,
    }


def cpu_baseline_prcc(seed, budget_s):
    """The reference's CPU validation path for C3's programs, one core:
    PageRankHost (f32 FIFO delta-push, epsilon 1e-6; restated in oracle/oracle.c)
    and a textbook CC label propagation (the reference has no CPU CC path), from
    scratch on G_0 of the C3 generator at scale 20 EF32 (a bounded sample: the
    full graph would take hours on one core)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    nv, row, col = sgraph.gen_rmat(20, C3["edge_factor"], seed=seed)
    ne = len(col)
    tp, tc = [], []
    t_start = time.perf_counter()
    while len(tp) < 2 or (time.perf_counter() - t_start < budget_s and len(tp) < 10):
        t0 = time.perf_counter()
        O.pagerank_host_f32(row, col)
        t1 = time.perf_counter()
        O.cc_label_prop(row, col)
        tc.append(time.perf_counter() - t1)
        tp.append(t1 - t0)
    p_med, c_med = float(np.median(tp)), float(np.median(tc))
    return {
        "value": round(2 * ne / (p_med + c_med) / 1e9, 6),
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"{len(tp)} x (one from-scratch PageRankHost (f32, epsilon 1e-6) + one CC label propagation) on G_0 "
                  f"of the C3 generator at scale 20 EF32 ({nv} vertices, {ne} edges), "
                  f"{time.perf_counter() - t_start:.1f} s in all; medians {p_med:.2f} s + {c_med:.3f} s",
        "host_cpus": os.cpu_count(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--c3-steps", type=int, default=2, help="timed C3 snapshots per path")
    ap.add_argument("--c3-warmup", type=int, default=0, help="untimed C3 snapshots per path after the static solve")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--batch", type=int, default=100000, help="inserts and deletes per snapshot")
    ap.add_argument("--algos", default=None, help="default sssp,bfs (incremental) or pr,cc (oom)")
    ap.add_argument("--seed", type=int, default=sgraph.DEFAULT_SEED)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no HIP-event kernel timing (no roofline)")
    ap.add_argument("--all-kernel-timing", action="store_true",
                    help="time every kernel class, not only the sweep (breakdown; the events slow the snapshot)")
    ap.add_argument("--wave-budget", type=int, default=0)
    ap.add_argument("--heavy-degree", type=int, default=0)
    ap.add_argument("--round-blocks", type=int, default=0)
    ap.add_argument("--rounds-per-sync", type=int, default=0, help="rounds enqueued per host check (0: engine default)")
    ap.add_argument("--graph-rounds", action="store_true", help="submit each batch of rounds as one HIP graph")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (shared-graph) measurement")
    ap.add_argument("--no-c4", action="store_true", help="default line: skip the C4 batched-queries key")
    ap.add_argument("--in-edges", action="store_true",
                    help="C2 engines keep the in-edge store: the recompute after deletions reads the in-lists of the "
                         "reset vertices instead of sweeping every live edge")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 256 MiB (5 times) to calibrate rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--traffic", type=float, default=None,
                    help="HBM bytes per launch of the roofline kernel measured with rocprofv3 FETCH_SIZE + WRITE_SIZE")
    ap.add_argument("--workload", default="c3", choices=["c3", "incremental", "oom", "serve", "batch"],
                    help="c3: the headline line (C3 + C2); incremental: C2 only; oom: out-of-memory path at any "
                         "scale; serve: request loop; batch: many queries")
    ap.add_argument("--edge-budget-gb", type=float, default=8.0, help="oom: device budget of the edge array")
    ap.add_argument("--cache", type=int, default=4, help="oom: -cache level")
    ap.add_argument("--segments", type=int, default=512, help="oom: SEGMENT count")
    ap.add_argument("--pma-density", type=float, default=0.0,
                    help="oom: PMA fill density at load (0: engine default); the host edge array is capacity x 4 B")
    ap.add_argument("--duration", type=float, default=20.0, help="serve: seconds of arrivals")
    ap.add_argument("--update-rate", type=float, default=200.0)
    ap.add_argument("--query-rate", type=float, default=500.0)
    ap.add_argument("--max-update", type=int, default=100000)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--clients", type=int, default=4)
    ap.add_argument("--queries", type=int, default=128, help="batch: BFS + SSSP queries")
    ap.add_argument("--plain-sample", type=int, default=16, help="batch: sources timed on single-source engines")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU "
              f"(every box has a single MI355X); run with --gpus 1", file=sys.stderr)
        return 2

    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 26, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": x.numel() * 4, "bytes_written": x.numel() * 4,
                                               "copies": 5}}), flush=True)
        return 0
    if args.algos is None:
        args.algos = "pr,cc" if args.workload == "oom" else "sssp,bfs"
    if args.workload == "oom":
        print(json.dumps(run_oom(args)), flush=True)
        return 0
    if args.workload == "serve":
        print(json.dumps(run_serve(args)), flush=True)
        return 0
    if args.workload == "batch":
        print(json.dumps(run_batch(args)), flush=True)
        return 0
    c2 = run_c2(args)
    if not args.no_cpu_baseline:
        log("[bench] C2 CPU baseline ...")
        c2["cpu_baseline"] = cpu_baseline(*c2.pop("_graph"), 12)
    else:
        c2.pop("_graph")
    if args.workload == "incremental":
        print(json.dumps(c2), flush=True)
        return 0
    t_c3 = time.perf_counter()
    c3 = run_c3(args)
    t_c3 = time.perf_counter() - t_c3
    res = c3["res"]
    lat_b = c3["lat_b"]

    def side(a, programs=1):
        r = res[a]
        st = r["stats"]
        return {"value": round(programs * sum(r["edges"]) / sum(r["lat"]) / 1e9, 4), "unit": "GTEPS",
                "latency_s": [round(x, 3) for x in r["lat"]], "load_s": round(r["load_s"], 1),
                "static_solve_s": round(r["init_s"], 1), "rounds_last": st["rounds"], "pull_rounds": st["pull_rounds"],
                "invalidated_last": st["invalidated"], "host_bytes": st["host_bytes"], "device_bytes": st["device_bytes"],
                "streamed_bytes": st["streamed_bytes"], "stream_ms": round(st["stream_ms"], 1),
                "cache_fill_bytes": st["cache_fill_bytes"], "cache_fill_ms": round(st["cache_fill_ms"], 1),
                "in_mirror_slots": st["in_mirror_slots"]}

    pulls = c3["pulls"]
    ms_pull = c3["ms_pull"]
    # memory traffic per dense pull launch, rocprofv3 FETCH_SIZE + WRITE_SIZE at C3 (committed summary)
    traffic_c3, traffic_c3_src = measured_traffic("sg::k_pr_pull", "k_pr_pull_c3_prcc_rmat27.pmc.txt")
    out = {
        "metric": "updated-edge GTEPS + p50/p99 snapshot latency, 1 MI355X, batched vs plain",
        "value": round(c3["batched"], 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": c3["steps"],
        "warmup": c3["warm"],
        "ms_per_step": round(float(np.mean(lat_b)) * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/f32",
        "data": "synthetic (seeded R-MAT graph and hub-biased update stream generated in-process)",
        "config": {
            "workload": "C3: PageRank + CC on R-MAT scale 27 EF32, edge arrays in pinned host memory under an 8 GiB "
                        "device edge budget (HBM copy of the in-store's leading slots), -cache 4, SEGMENT 512, "
                        "1,000,000 inserts (80% from the top-1% sources) + 1,000,000 deletes per snapshot; "
                        "batched = both programs on one engine (SG_ALGO_PR_CC)",
            "scale": C3["scale"], "edge_factor": C3["edge_factor"], "nv": c3["nv"], "ne0": c3["ne0"],
            "batch_inserts": C3["batch"], "batch_deletes": C3["batch"], "pma_density": C3["pma_density"],
            "pr_epsilon": 1e-6, "programs_per_snapshot": 2, "steps_note":
                "--steps/--warmup set the C2 line; C3 times --c3-steps snapshots per path (tens of seconds each)",
        },
        "latency_p50_ms": pct(lat_b, 50),
        "latency_p99_ms": pct(lat_b, 99),
        "batched": {**side("prcc", 2), "vs_plain": round(c3["batched"] / c3["plain"], 3)},
        "plain": {"value": round(c3["plain"], 4), "unit": "GTEPS", "pr": side("pr"), "cc": side("cc"),
                  "latency_p50_ms": pct([a + b for a, b in zip(res["pr"]["lat"], res["cc"]["lat"])], 50),
                  "latency_p99_ms": pct([a + b for a, b in zip(res["pr"]["lat"], res["cc"]["lat"])], 99)},
        "roofline": {
            "kernel": "k_pr_pull + k_pr_pull_fix (PageRank pull rounds over the in-edge store: 8 B per in-edge + "
                      "24 B per vertex); in-lists read from a packed HBM copy, the rest from a packed host copy",
            "bound": "hbm",
            "achieved": round(c3["achieved"], 2) if c3["achieved"] else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(c3["achieved"] / HBM_PEAK_GBS, 4) if c3["achieved"] else None,
            "traffic": traffic_c3,
            "traffic_source": traffic_c3_src,
            "launches": pulls,
            "avg_launch_us": round(ms_pull * 1e3 / pulls, 1) if pulls else None,
            "algorithmic_bytes_per_launch": round(c3["b_pull"] / pulls) if pulls else None,
        },
        "host_link": {
            "pull_rounds_gbs": round(c3["host_per_pull"] * pulls / (ms_pull * 1e-3) / 1e9, 2) if ms_pull else None,
            "pull_rounds_what": "packed bytes of the live in-edges outside the HBM copy per pull round x pull rounds "
                                "/ pull time",
            "cc_streamed_sweep_gbs": round(res["cc"]["stats"]["streamed_bytes"] / (res["cc"]["stats"]["stream_ms"] * 1e-3)
                                           / 1e9, 2) if res["cc"]["stats"]["stream_ms"] else None,
            "cache_fill_gbs": round(res["prcc"]["stats"]["cache_fill_bytes"] / (res["prcc"]["stats"]["cache_fill_ms"] * 1e-3)
                                    / 1e9, 2) if res["prcc"]["stats"]["cache_fill_ms"] else None,
            "peak_gbs": 64.0,
        },
        "c3_wall_s": round(t_c3, 1),
        "c2": c2,
    }
    if not args.no_c4:
        # C4: 64 BFS + 64 SSSP queries on R-MAT scale 24 EF16 under batches of
        # log-uniform size in [1e3, 1e6]: batched engine vs single-source engines
        a4 = argparse.Namespace(**vars(args))
        a4.scale, a4.edge_factor, a4.queries, a4.max_update, a4.steps, a4.warmup = 24, 16, 128, 1000000, 4, 1
        log("[bench] C4 batched queries ...")
        t_c4 = time.perf_counter()
        c4 = run_batch(a4)
        c4["wall_s"] = round(time.perf_counter() - t_c4, 1)
        out["c4"] = {k: c4[k] for k in ("metric", "value", "unit", "vs_plain", "plain", "batched_reached_mean",
                                         "latency_p50_ms", "latency_p99_ms", "config", "steps", "warmup", "wall_s")}
    if not args.no_cpu_baseline:
        log("[bench] C3 CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline_prcc(args.seed, 15)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
