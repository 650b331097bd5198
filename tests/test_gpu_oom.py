"""Out-of-memory path: the edge array in pinned host memory (forced with
SG_FLAG_HOST_EDGES or by an edge budget smaller than the array), the boundary
sweep streamed segment by segment through HBM staging buffers, and the HBM
edge cache. Results must equal the oracle's (bit-exact for BFS/SSSP/CC,
PageRank within the tolerances of the resident tests) after every snapshot,
with and without the cache; the SEGMENT cuts equal the oracle's restatement
of Engine::LoadGraph's greedy split (framework.cuh:310-362)."""
import numpy as np
import pytest

import oracle as O
import sgraph
from test_gpu_engine import check_pma, check_values

pytestmark = pytest.mark.gpu


def run_host(algo, scale, ef, nsnap, nupd, segments, cache_level, seed=31, **cfg):
    nv, row, col = sgraph.gen_rmat(scale, ef, seed=seed)
    ins, dele = sgraph.gen_updates(row, col, scale, nsnap, nupd, nupd, seed=seed + 1)
    print(f"[gpu] host-edges {algo} nv={nv} ne={len(col)} seg={segments} cache={cache_level} {cfg}", flush=True)
    e = sgraph.Engine(algo, weight_mode="seeded", weight_seed=3, segments=segments, cache_level=cache_level, **cfg)
    try:
        e.load_csr(row, col)
        st = e.stats()
        assert st["host_edges"] == 1
        segs = e.gather(sgraph.Q_SEGMENTS)
        off = e.gather(sgraph.Q_OFFSETS)
        exp = O.segment_partition(off, segments)
        assert segs[:len(exp)].tolist() == exp.tolist()
        assert np.all(segs[len(exp):] == nv)
        check_pma(e, row, col, "load")
        e.init(0 if algo in ("bfs", "sssp") else None)
        check_values(algo, e, row, col, 0, "seeded", 3, "G0")
        r, c = row, col
        for t in range(nsnap):
            e.apply(ins[t], dele[t])
            r, c = O.apply_updates(r, c, ins[t], dele[t])
            check_pma(e, r, c, f"G{t + 1}")
            check_values(algo, e, r, c, 0, "seeded", 3, f"G{t + 1}")
        return e.stats()
    finally:
        e.close()


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc", "pr"])
def test_host_edges_streamed_with_cache(algo):
    st = run_host(algo, 12, 8, 3, 500, segments=7, cache_level=1, flags=sgraph.FLAG_HOST_EDGES)
    # traversals stream the boundary sweep and cache hot out-lists (refills
    # over the link counted apart); PageRank's pull rounds read its in-edge
    # store in place (unless SG_FLAG_PULL_STREAM), and the -cache budget holds
    # the store's HBM copy instead of out-lists
    if algo == "pr":
        assert st["cache_slots"] == 0 and st["in_mirror_live"] == st["ne"]
    else:
        assert st["cache_slots"] > 0 and st["cache_fill_bytes"] > 0
        assert st["streamed_bytes"] > 0 and st["stream_ms"] > 0


@pytest.mark.parametrize("algo", ["sssp", "cc"])
def test_host_edges_without_cache(algo):
    st = run_host(algo, 11, 8, 2, 300, segments=3, cache_level=0, flags=sgraph.FLAG_HOST_EDGES)
    assert st["cache_slots"] == 0 and st["edges_from_cache"] == 0


def test_edge_budget_moves_edges_to_host_and_forces_rebalance():
    # a budget far below the edge array; a dense PMA so the batches rebalance and resize
    st = run_host("bfs", 11, 8, 3, 1500, segments=16, cache_level=2, edge_budget_gb=1e-6, pma_density=0.9)
    assert st["rebalances"] + st["resizes"] > 0


def test_cache_serves_edges():
    # PageRank with push rounds only (no in-edge store): its rounds walk
    # out-lists, and the hot ones come from the HBM cache
    nv, row, col = sgraph.gen_rmat(12, 16, seed=5)
    ins, dele = sgraph.gen_updates(row, col, 12, 3, 200, 200, seed=6)
    e = sgraph.Engine("pr", segments=4, cache_level=1, flags=sgraph.FLAG_HOST_EDGES, pull_divisor=sgraph.PULL_NEVER)
    e.load_csr(row, col)
    e.init()
    assert e.stats()["cached_vertices"] > 0
    r, c = row, col
    for t in range(3):
        e.apply(ins[t], dele[t])
        r, c = O.apply_updates(r, c, ins[t], dele[t])
    check_values("pr", e, r, c, 0, "hash128", 0, "G3")
    st = e.stats()
    assert st["edges_from_cache"] > 0
    e.close()


def test_device_budget_refuses_with_ebudget():
    # device_budget_gb caps the engine's accounted device memory: a load that
    # needs more fails with SG_EBUDGET and a message naming the budget, and a
    # fresh engine with room for the load works
    nv, row, col = sgraph.gen_rmat(14, 16, seed=61)
    e = sgraph.Engine("sssp", device_budget_gb=1e-4)  # 100 KiB: below the vertex state alone
    with pytest.raises(sgraph.SgError) as x:
        e.load_csr(row, col)
    assert x.value.status == sgraph.SG_EBUDGET and "budget" in str(x.value)
    e.close()
    f = sgraph.Engine("sssp", device_budget_gb=1.0)
    f.load_csr(row, col)
    f.init(0)
    assert np.array_equal(f.values(), O.sssp(row, col, 0))
    assert f.stats()["device_bytes"] <= 2**30
    f.close()
