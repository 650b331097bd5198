"""Parity at the headline workload's full size: R-MAT scale 22, edge factor
16 (65.2M edges), SSSP with seeded [1, 63] weights and BFS from source 0,
snapshots of 100,000 fresh inserts + 100,000 deletes. Every distance and
level is compared with the oracle (Dijkstra / BFS in C on the host) after
the static solve and after each snapshot, bit-exact; the batched engine (both
queries on one graph) must give the same vectors."""
import numpy as np
import pytest

import oracle as O
import sgraph

pytestmark = pytest.mark.gpu


def test_scale22_sssp_bfs_bit_exact():
    nv, row, col = sgraph.gen_rmat(22, 16, seed=sgraph.DEFAULT_SEED)
    ins, dele = sgraph.gen_updates(row, col, 22, 2, 100000, 100000, seed=sgraph.DEFAULT_SEED + 1)
    wseed = sgraph.DEFAULT_SEED + 2
    print(f"[gpu] full size nv={nv} ne={len(col)}", flush=True)
    engines = {a: sgraph.Engine(a, weight_mode="seeded", weight_seed=wseed) for a in ("sssp", "bfs")}
    multi = sgraph.Engine("multi", weight_mode="seeded", weight_seed=wseed)
    try:
        for e in engines.values():
            e.load_csr(row, col)
            e.init(0)
        multi.load_csr(row, col)
        multi.init_queries([0, 0], ["sssp", "bfs"])
        r, c = row, col
        for t in range(3):
            if t:
                for e in engines.values():
                    e.apply(ins[t - 1], dele[t - 1])
                multi.apply(ins[t - 1], dele[t - 1])
                r, c = O.apply_updates(r, c, ins[t - 1], dele[t - 1])
            exp_s = O.sssp(r, c, 0, "seeded", wseed)
            exp_b = O.bfs(r, c, 0)
            got_s, got_b = engines["sssp"].values(), engines["bfs"].values()
            assert np.array_equal(got_s, exp_s), f"G{t} sssp: {np.count_nonzero(got_s != exp_s)} mismatches"
            assert np.array_equal(got_b, exp_b), f"G{t} bfs: {np.count_nonzero(got_b != exp_b)} mismatches"
            assert np.array_equal(multi.query(0), got_s) and np.array_equal(multi.query(1), got_b), f"G{t} batched"
            print(f"[gpu] G{t} ok", flush=True)
    finally:
        for e in engines.values():
            e.close()
        multi.close()
