"""Headline benchmark: updated-edge GTEPS of incremental SSSP + BFS on one MI355X.

Workload (the plain per-snapshot path on the resident configuration): seeded
R-MAT, 2^22 vertices, edge factor 16 (a/b/c/d = 0.57/0.19/0.19/0.05, self-loops
and duplicates dropped), SSSP with seeded weights in [1, 63] and BFS, both
from source 0; every step is one snapshot of 100,000 fresh inserts + 100,000
deletes sampled from the current edges, applied to the SSSP engine and to the
BFS engine and recomputed to convergence.

  updated-edge GTEPS = sum_t (E_t per query answered) / sum_t T_t / 1e9
  T_t = wall time from hand-off of the device-resident batch to converged
        device state (PMA update + invalidation + recompute)

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


def cpu_baseline(row, col, seed, budget_s):
    """The reference's CPU validation path (SSSPHostNaive + BFSHost, restated in
    oracle/oracle.c) from scratch on the snapshot graph, one core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.lib()
    ne = len(col)
    t0 = time.perf_counter()
    O.sssp_fifo(row, col, 0, "seeded", seed)
    t1 = time.perf_counter()
    O.bfs(row, col, 0)
    t2 = time.perf_counter()
    return {
        "value": round(2 * ne / (t2 - t0) / 1e9, 6),
        "unit": "GTEPS",
        "cores": 1,
        "kind": "port",
        "sample": f"one from-scratch SSSP (FIFO Bellman-Ford, SSSPHostNaive) + one BFS (BFSHost) on G_0 "
                  f"({len(row) - 1} vertices, {ne} edges); {t1 - t0:.2f} s + {t2 - t1:.2f} s",
        "host_cpus": os.cpu_count(),
    }


def run_batched(args, row, col, ins, dele, wseed, algos, plain_value):
    """The same queries answered together on one shared graph (one PMA update,
    one boundary sweep and one launch per round for all of them)."""
    nsnap = args.warmup + args.steps
    e = sgraph.Engine("multi", weight_mode="seeded", weight_seed=wseed)
    t = time.perf_counter()
    e.load_csr(row, col)
    e.init_queries([0] * len(algos), algos)
    log(f"[bench] batched {'+'.join(algos)}: load+static solve {time.perf_counter() - t:.2f}s")
    dev = [(e.to_device(ins[s]), e.to_device(dele[s])) for s in range(nsnap)]
    for s in range(args.warmup):
        e.apply_device(dev[s][0], args.batch, dev[s][1], args.batch)
    e.synchronize()
    lat, num = [], 0
    for s in range(args.warmup, nsnap):
        t0 = time.perf_counter()
        e.apply_device(dev[s][0], args.batch, dev[s][1], args.batch)
        lat.append(time.perf_counter() - t0)
        num += e.stats()["ne"] * len(algos)
    value = num / sum(lat) / 1e9
    for di, dd in dev:
        e.device_free(di)
        e.device_free(dd)
    e.close()
    lat_ms = np.array(lat) * 1e3
    return {
        "value": round(value, 4), "unit": "GTEPS", "vs_plain": round(value / plain_value, 3),
        "latency_p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "latency_p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "what": f"{len(algos)} queries ({'+'.join(algos)} from source 0) on one shared graph: one PMA update, one "
                f"sweep and one launch per round per snapshot (sg_init_queries)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--batch", type=int, default=100000, help="inserts and deletes per snapshot")
    ap.add_argument("--algos", default="sssp,bfs")
    ap.add_argument("--seed", type=int, default=sgraph.DEFAULT_SEED)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="do not time the relax kernel with HIP events")
    ap.add_argument("--wave-budget", type=int, default=0)
    ap.add_argument("--heavy-degree", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (shared-graph) measurement")
    ap.add_argument("--batched-only", action="store_true", help="measure only the batched path (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU "
              f"(every box has a single MI355X); run with --gpus 1", file=sys.stderr)
        return 2

    algos = args.algos.split(",")
    nsnap = args.warmup + args.steps
    t = time.perf_counter()
    nv, row, col = sgraph.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    ins, dele = sgraph.gen_updates(row, col, args.scale, nsnap, args.batch, args.batch, seed=args.seed + 1)
    log(f"[bench] graph nv={nv} ne={len(col)} snapshots={nsnap} batch={args.batch} gen {time.perf_counter() - t:.1f}s")
    wseed = args.seed + 2

    if args.batched_only:
        b = run_batched(args, row, col, ins, dele, wseed, algos, 1.0)
        print(json.dumps({"batched": b}), flush=True)
        return 0

    engines = {}
    dev_batches = {}
    for a in algos:
        flags = 0 if args.no_timing else sgraph.FLAG_KERNEL_TIMING
        e = sgraph.Engine(a, weight_mode="seeded", weight_seed=wseed, flags=flags, wave_budget=args.wave_budget,
                          heavy_degree=args.heavy_degree)
        t = time.perf_counter()
        e.load_csr(row, col)
        e.init(0)
        st = e.stats()
        log(f"[bench] {a}: load+static solve {time.perf_counter() - t:.2f}s rounds={st['rounds']}")
        # every snapshot's pairs resident in HBM before the timed region
        dev_batches[a] = [(e.to_device(ins[s]), e.to_device(dele[s])) for s in range(nsnap)]
        engines[a] = e

    ne_t = len(col)
    edges_per_step = []
    lat = {a: [] for a in algos}
    for s in range(args.warmup):
        for a in algos:
            di, dd = dev_batches[a][s]
            engines[a].apply_device(di, args.batch, dd, args.batch)
    for a in algos:
        engines[a].reset_stats()
        engines[a].synchronize()
    t_start = time.perf_counter()
    for s in range(args.warmup, nsnap):
        e_t = 0
        for a in algos:
            e = engines[a]
            di, dd = dev_batches[a][s]
            t0 = time.perf_counter()
            e.apply_device(di, args.batch, dd, args.batch)  # returns after convergence (device synchronised)
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
    relax_ms = sum(stats[a]["relax_ms"] for a in algos)
    relax_bytes = sum(stats[a]["relax_bytes"] for a in algos)
    launches = sum(stats[a]["relax_launches"] for a in algos)
    achieved = relax_bytes / (relax_ms * 1e-3) / 1e9 if relax_ms > 0 else None
    out = {
        "metric": "updated-edge GTEPS (incremental SSSP+BFS per snapshot, 1 MI355X, plain per-snapshot path)",
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
            "kernel": "k_round (convergence round: light items + spilled chunks of the previous round)",
            "bound": "hbm",
            "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": None,
            "launches": launches,
            "avg_launch_us": round(relax_ms * 1e3 / launches, 2) if launches else None,
            "algorithmic_bytes_per_launch": round(relax_bytes / launches) if launches else None,
        },
        "engine": {a: {k: stats[a][k] for k in ("rounds", "invalidated", "rebalances", "resizes", "edges_relaxed",
                                               "vertices_processed", "device_bytes")} for a in algos},
    }
    if not args.no_batched:
        out["batched"] = run_batched(args, row, col, ins, dele, wseed, algos, value)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(row, col, wseed, 30)
    for a in algos:
        for di, dd in dev_batches[a]:
            engines[a].device_free(di)
            engines[a].device_free(dd)
        engines[a].close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
