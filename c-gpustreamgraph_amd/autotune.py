"""Autotuner for the engine's launch shapes and out-of-memory split.

No counterpart in the reference (its block size, SEGMENT count, -cache and
stream count are command-line constants, app_skeleton.h:42-100). Given a
graph and a few update batches of the workload, every candidate of the search
space runs the same snapshots on a fresh engine; the configuration with the
lowest total snapshot time wins. Candidates only change how work is
scheduled (wave budget, heavy-list threshold, rounds per host check, round
batches as HIP graphs, SEGMENT count, cache level), never what is computed, so
every candidate gives the same results.
"""
import itertools
import time

import sgraph

# knobs of the resident path
RESIDENT_SPACE = {
    "wave_budget": [512, 1024, 2048],
    "heavy_degree": [2048, 8192],
    "rounds_per_sync": [0, 2, 4, 8],
    "round_blocks": [1024, 2048],
    "graph_rounds": [False, True],  # batches of rounds as HIP graphs (SG_FLAG_GRAPH_ROUNDS)
}
# knobs of the out-of-memory path
HOST_SPACE = {
    "segments": [64, 256, 512],
    "cache_level": [0, 1, 2, 4],
}


def engine_config(cfg):
    """Engine keyword arguments of a search-space point (graph_rounds becomes a flag)."""
    cfg = dict(cfg)
    if cfg.pop("graph_rounds", False):
        cfg["flags"] = cfg.get("flags", 0) | sgraph.FLAG_GRAPH_ROUNDS
    return cfg


def _run(algo, row, col, ins, dele, source, cfg):
    e = sgraph.Engine(algo, **engine_config(cfg))
    try:
        e.load_csr(row, col)
        if algo == "multi":
            e.init_queries(source, ["sssp"] * len(source))
        else:
            e.init(source if algo in ("bfs", "sssp") else None)
        e.synchronize()
        t0 = time.perf_counter()
        for i, d in zip(ins, dele):
            e.apply(i, d)
        e.synchronize()
        return time.perf_counter() - t0
    finally:
        e.close()


def autotune(algo, row, col, ins, dele, source=0, space=None, repeats=1, **base):
    """Time every point of `space` (dict of knob -> values; by default the
    resident knobs, plus the out-of-memory knobs when base asks for host
    edges) on the given snapshots. Returns (engine keyword arguments of the
    fastest point, [(point, seconds)])."""
    if space is None:
        space = dict(RESIDENT_SPACE)
        if base.get("flags", 0) & sgraph.FLAG_HOST_EDGES or base.get("edge_budget_gb", 0) > 0:
            space.update(HOST_SPACE)
    keys = sorted(space)
    table = []
    for values in itertools.product(*(space[k] for k in keys)):
        cfg = dict(base)
        cfg.update(zip(keys, values))
        t = min(_run(algo, row, col, ins, dele, source, cfg) for _ in range(max(1, repeats)))
        table.append((cfg, t))
    best = min(table, key=lambda x: x[1])[0]
    return engine_config(best), table


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="pick launch shapes for a workload")
    ap.add_argument("--algo", default="sssp")
    ap.add_argument("--scale", type=int, default=18)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--batch", type=int, default=10000)
    ap.add_argument("--snapshots", type=int, default=3)
    ap.add_argument("--seed", type=int, default=sgraph.DEFAULT_SEED)
    args = ap.parse_args()
    nv, row, col = sgraph.gen_rmat(args.scale, args.edge_factor, seed=args.seed)
    ins, dele = sgraph.gen_updates(row, col, args.scale, args.snapshots, args.batch, args.batch, seed=args.seed + 1)
    best, table = autotune(args.algo, row, col, ins, dele)
    print(json.dumps({"best": best, "candidates": [{"cfg": c, "s": round(t, 6)} for c, t in table]}))


if __name__ == "__main__":
    main()
