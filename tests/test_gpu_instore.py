"""Traversal recompute seeded from the in-edge store, and the PageRank + CC
engine that shares one forward store and one in-edge store.

* SSSP / BFS / CC with SG_FLAG_IN_EDGES: after every snapshot the results are
  bit-exact against the oracle (the reset vertices take their best value from
  their own in-lists instead of a sweep over every edge), resident and with
  both stores in pinned host memory, including in-lists long enough to be cut
  into chunks (small wave budget).
* PR + CC engine: per snapshot, CC labels bit-exact against the oracle's
  min-ancestor labels and PageRank within the PR bounds of the FP64 fixed
  point; both equal to what separate single-program engines give.
"""
import numpy as np
import pytest

import oracle as O
import sgraph
from test_gpu_engine import check_values
from test_gpu_pull import check_in_store, pr_errors

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def trav_stream(algo, scale, ef, nsnap, nupd, seed, check_store=False, **cfg):
    nv, row, col = sgraph.gen_rmat(scale, ef, seed=seed)
    ins, dele = sgraph.gen_updates(row, col, scale, nsnap, nupd, nupd, seed=seed + 1, hub_frac=0.5, hub_top=0.01)
    flags = sgraph.FLAG_IN_EDGES | cfg.pop("flags", 0)
    say(f"[gpu] in-edges {algo} nv={nv} ne={len(col)} flags={flags} {cfg}")
    e = sgraph.Engine(algo, weight_mode="seeded", weight_seed=7, flags=flags, **cfg)
    try:
        e.load_csr(row, col)
        e.init(0 if algo in ("bfs", "sssp") else None)
        check_values(algo, e, row, col, 0, "seeded", 7, "G0")
        r, c = row, col
        inval = 0
        for t in range(nsnap):
            e.apply(ins[t], dele[t])
            r, c = O.apply_updates(r, c, ins[t], dele[t])
            if check_store:
                check_in_store(e, r, c, f"G{t + 1}")
            check_values(algo, e, r, c, 0, "seeded", 7, f"G{t + 1}")
            inval += e.stats()["invalidated"]
        st = e.stats()
        assert st["in_capacity"] > 0 and inval > 0
        return st
    finally:
        e.close()


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_in_edges_seed_resident(algo):
    trav_stream(algo, 14, 16, 3, 4000, seed=81, check_store=True)


@pytest.mark.parametrize("algo", ["sssp", "cc"])
def test_in_edges_seed_long_in_lists_cut_into_chunks(algo):
    # a small wave budget sends most in-lists to the chunked tier: the pull
    # seed's spilled descriptors index the in-store and are walked before round R
    trav_stream(algo, 13, 32, 3, 3000, seed=83, wave_budget=64, heavy_degree=1024)


@pytest.mark.parametrize("algo", ["sssp", "cc"])
def test_in_edges_seed_host_edges(algo):
    st = trav_stream(algo, 12, 16, 3, 2000, seed=85, flags=sgraph.FLAG_HOST_EDGES, segments=7, cache_level=1)
    assert st["host_edges"] == 1


def prcc_stream(scale, ef, nsnap, nupd, seed, **cfg):
    nv, row, col = sgraph.gen_rmat(scale, ef, seed=seed)
    ins, dele = sgraph.gen_updates(row, col, scale, nsnap, nupd, nupd, seed=seed + 1, hub_frac=0.8, hub_top=0.01)
    say(f"[gpu] prcc nv={nv} ne={len(col)} {cfg}")
    e = sgraph.Engine("prcc", **cfg)
    single = {a: sgraph.Engine(a, **cfg) for a in ("pr", "cc")}
    try:
        for x in [e] + list(single.values()):
            x.load_csr(row, col)
            x.init()
        r, c = row, col
        for t in range(nsnap + 1):
            if t:
                for x in [e] + list(single.values()):
                    x.apply(ins[t - 1], dele[t - 1])
                r, c = O.apply_updates(r, c, ins[t - 1], dele[t - 1])
            labels = e.query(1)
            assert np.array_equal(labels, O.cc(r, c)), f"G{t}: CC labels differ from the oracle"
            assert np.array_equal(labels, single["cc"].values()), f"G{t}: CC differs from the CC engine"
            assert np.array_equal(e.gather_at(np.arange(0, nv, 7), qi=1), labels[::7])
            ranks = e.query(0)
            rel, nrm, l1 = pr_errors(ranks, O.pagerank(r, c))
            rel1, _, _ = pr_errors(single["pr"].values(), O.pagerank(r, c))
            say(f"[gpu] prcc G{t} PR rel={rel:.2e} norm={nrm:.2e} L1={l1:.2e} (PR engine rel={rel1:.2e}) "
                f"invalidated={e.stats()['invalidated']}")
            assert rel <= 5e-5 and nrm <= 1e-6
        return e.stats()
    finally:
        e.close()
        for x in single.values():
            x.close()


def test_prcc_resident():
    prcc_stream(13, 16, 3, 3000, seed=87)


def test_prcc_host_edges():
    st = prcc_stream(12, 16, 3, 2000, seed=89, flags=sgraph.FLAG_HOST_EDGES, segments=7, cache_level=0)
    assert st["host_edges"] == 1 and st["in_capacity"] > 0 and st["in_mirror_live"] == 0


def test_prcc_host_edges_with_hbm_mirror():
    # the edge budget (below the edge array) holds an HBM copy of part of the
    # in-store: CC's in-list seeds and PageRank's pulls read those slots there
    st = prcc_stream(12, 16, 3, 2000, seed=93, edge_budget_gb=2e-5, segments=7, cache_level=0)
    assert st["host_edges"] == 1 and 0 < st["in_mirror_live"] < st["ne"]


def test_prcc_deferred_pagerank():
    # PageRank deferred, CC recomputed every snapshot
    nv, row, col = sgraph.gen_rmat(12, 16, seed=91)
    ins, dele = sgraph.gen_updates(row, col, 12, 3, 1500, 1500, seed=92)
    e = sgraph.Engine("prcc", flags=sgraph.FLAG_DEFER_CONVERGE)
    e.load_csr(row, col)
    e.init()
    r, c = row, col
    for t in range(3):
        e.apply(ins[t], dele[t])
        r, c = O.apply_updates(r, c, ins[t], dele[t])
        assert np.array_equal(e.query(1), O.cc(r, c))
    rel, nrm, _ = pr_errors(e.query(0), O.pagerank(r, c))  # converges the deferred snapshots
    e.close()
    assert rel <= 5e-5 and nrm <= 1e-6
