"""PageRank pull rounds over the in-edge store.

* The in-edge store holds, after every snapshot, exactly the transpose of the
  update model's graph (per vertex, the multiset of in-neighbours), with the
  PMA invariants.
* PageRank with push rounds only, with every round pulled, and with the
  default direction choice each meets the FP64 fixed point within the bounds
  of the other PR tests (5e-5 relative on raw ranks, 1e-6 max-abs on
  L1-normalised ranks), resident and with the store streamed from host memory.
* C1 at its full size: R-MAT scale 16 EF16, 2 snapshots of 10,000 inserts +
  10,000 deletes, checked against the FP64 oracle with the L1 error printed.
"""
import numpy as np
import pytest

import oracle as O
import sgraph

pytestmark = pytest.mark.gpu

PR_REL_TOL = 5e-5
PR_NORM_TOL = 1e-6


def say(*a):
    print(*a, flush=True)


def pr_errors(got, exp):
    got = got.astype(np.float64)
    rel = float(np.max(np.abs(got - exp) / exp))
    gn, en = got / got.sum(), exp / exp.sum()
    return rel, float(np.max(np.abs(gn - en))), float(np.abs(gn - en).sum())


def transpose_lists(row, col):
    nv = len(row) - 1
    src = np.repeat(np.arange(nv, dtype=np.uint32), np.diff(row.astype(np.int64)))
    order = np.lexsort((src, col))
    trow = np.zeros(nv + 1, dtype=np.uint64)
    np.cumsum(np.bincount(col, minlength=nv).astype(np.uint64), out=trow[1:])
    return trow, src[order]


def check_in_store(eng, row, col, tag):
    r, c, off, deg, slots = eng.in_adjacency()
    trow, tcol = transpose_lists(row, col)
    assert np.array_equal(r, trow), f"{tag}: in-degrees differ"
    # per-vertex multisets: sort within each list
    nv = len(r) - 1
    key_got = np.repeat(np.arange(nv, dtype=np.uint64), np.diff(r.astype(np.int64))) << np.uint64(32) | c.astype(np.uint64)
    key_exp = np.repeat(np.arange(nv, dtype=np.uint64), np.diff(trow.astype(np.int64))) << np.uint64(32) | tcol.astype(np.uint64)
    assert np.array_equal(np.sort(key_got), np.sort(key_exp)), f"{tag}: in-lists differ"
    cap = len(slots)
    assert off[-1] == cap and np.all(off[:-1] + deg <= off[1:]), f"{tag}: in-store ranges"
    assert int((slots != 0xFFFFFFFF).sum()) == int(r[-1]), f"{tag}: live slots outside the lists"


def pr_stream(row, col, snaps, check_store=False, **cfg):
    eng = sgraph.Engine("pr", **cfg)
    try:
        eng.load_csr(row, col)
        if check_store:
            check_in_store(eng, row, col, "load")
        eng.init()
        rel, nrm, l1 = pr_errors(eng.values(), O.pagerank(row, col))
        say(f"[gpu] pr G0 cfg={cfg} rel={rel:.2e} max-abs(norm)={nrm:.2e} L1(norm)={l1:.2e} "
            f"pull_rounds={eng.stats()['pull_rounds']}")
        assert rel <= PR_REL_TOL and nrm <= PR_NORM_TOL
        r, c = row, col
        for t, (ins, dele) in enumerate(snaps):
            eng.apply(ins, dele)
            r, c = O.apply_updates(r, c, ins, dele)
            if check_store:
                check_in_store(eng, r, c, f"G{t + 1}")
            rel, nrm, l1 = pr_errors(eng.values(), O.pagerank(r, c))
            st = eng.stats()
            say(f"[gpu] pr G{t + 1} rel={rel:.2e} max-abs(norm)={nrm:.2e} L1(norm)={l1:.2e} rounds={st['rounds']} "
                f"pull_rounds={st['pull_rounds']} residual_max={st['pr_residual_max']:.2e}")
            assert rel <= PR_REL_TOL, f"G{t + 1}: rel {rel}"
            assert nrm <= PR_NORM_TOL, f"G{t + 1}: normalised {nrm}"
            assert st["pr_residual_max"] <= cfg.get("pr_epsilon", 1e-6) * 1.0001
        return eng.stats()
    finally:
        eng.close()


def test_in_store_is_the_transpose():
    nv, row, col = sgraph.gen_rmat(12, 8, seed=51)
    ins, dele = sgraph.gen_updates(row, col, 12, 4, 1500, 1500, seed=52, hub_frac=0.8, hub_top=0.01)
    pr_stream(row, col, list(zip(ins, dele)), check_store=True)


def test_in_store_rebalance_and_resize():
    # a dense in-store whose hub targets gain many in-edges: windows re-spread and the capacity doubles
    nv, row, col = sgraph.gen_rmat(11, 8, seed=53)
    ins, dele = sgraph.gen_updates(row, col, 11, 4, 3000, 200, seed=54, hub_frac=0.9, hub_top=0.01)
    pr_stream(row, col, list(zip(ins, dele)), check_store=True, pma_density=0.9)


@pytest.mark.parametrize("mode", ["never", "always", "default"])
def test_pr_pull_modes(mode):
    div = {"never": sgraph.PULL_NEVER, "always": sgraph.PULL_ALWAYS, "default": 0}[mode]
    nv, row, col = sgraph.gen_rmat(14, 16, seed=55)
    ins, dele = sgraph.gen_updates(row, col, 14, 3, 3000, 3000, seed=56, hub_frac=0.8, hub_top=0.01)
    st = pr_stream(row, col, list(zip(ins, dele)), pull_divisor=div)
    if mode == "never":
        assert st["pull_rounds"] == 0 and st["in_capacity"] == 0
    else:
        assert st["pull_rounds"] > 0


@pytest.mark.parametrize("mode", ["always", "default"])
@pytest.mark.parametrize("stream", [True, False])
def test_pr_pull_host_edges(mode, stream):
    # forward and in-edge stores in pinned host memory: pull rounds stream the
    # in-store's SEGMENT slot ranges through HBM staging buffers, or read its
    # live slots in place over the host link
    div = {"always": sgraph.PULL_ALWAYS, "default": 0}[mode]
    nv, row, col = sgraph.gen_rmat(12, 16, seed=57)
    ins, dele = sgraph.gen_updates(row, col, 12, 3, 2000, 2000, seed=58, hub_frac=0.8, hub_top=0.01)
    flags = sgraph.FLAG_HOST_EDGES | (sgraph.FLAG_PULL_STREAM if stream else 0)
    # cache_level 0: no HBM copy of the in-store, every pull reads host memory
    st = pr_stream(row, col, list(zip(ins, dele)), pull_divisor=div, flags=flags, segments=7, cache_level=0)
    assert st["host_edges"] == 1 and st["pull_rounds"] > 0 and st["in_mirror_live"] == 0
    assert (st["streamed_bytes"] > 0) == stream


@pytest.mark.parametrize("budget,cache", [(2e-5, 0), (1.0, 0), (2e-5, 1)])
def test_pr_pull_host_edges_with_hbm_mirror(budget, cache):
    # both stores in pinned host memory; the edge budget (plus, for PageRank,
    # the -cache budget) holds an HBM copy of the in-store's leading live
    # edges, ids packed at ceil(log2 nv) bits - part of them, or all -
    # rebuilt after every batch; the rest come from a packed host copy
    nv, row, col = sgraph.gen_rmat(12, 16, seed=65)
    ins, dele = sgraph.gen_updates(row, col, 12, 3, 2000, 2000, seed=66, hub_frac=0.8, hub_top=0.01)
    st = pr_stream(row, col, list(zip(ins, dele)), check_store=True, edge_budget_gb=budget, cache_level=cache,
                   flags=sgraph.FLAG_HOST_EDGES, segments=7, pull_divisor=sgraph.PULL_ALWAYS)
    assert st["host_edges"] == 1 and st["pull_rounds"] > 0
    ids = (int(budget * 2**30) // 4 - 1) * 32 // 12  # 12-bit ids at 2^12 vertices
    partial = ids < st["ne"] and cache == 0
    assert 0 < st["in_mirror_live"] and (st["in_mirror_live"] < st["ne"]) == partial


def test_in_store_built_on_the_device():
    # from 2^24 edges the in-lists are sorted on the device: the same multisets
    nv, row, col = sgraph.gen_rmat_device(21, 16)
    assert len(col) >= (1 << 24)
    ins, dele = sgraph.gen_updates(row, col, 21, 1, 20000, 20000, seed=63)
    eng = sgraph.Engine("cc", flags=sgraph.FLAG_IN_EDGES)
    try:
        eng.load_csr(row, col)
        check_in_store(eng, row, col, "load")
        eng.init()
        eng.apply(ins[0], dele[0])
        r, c = O.apply_updates(row, col, ins[0], dele[0])
        check_in_store(eng, r, c, "G1")
        assert np.array_equal(eng.values(), O.cc(r, c))
    finally:
        eng.close()


def test_pr_deferred_with_pull():
    nv, row, col = sgraph.gen_rmat(12, 16, seed=59)
    ins, dele = sgraph.gen_updates(row, col, 12, 3, 2000, 2000, seed=60)
    eng = sgraph.Engine("pr", flags=sgraph.FLAG_DEFER_CONVERGE)
    eng.load_csr(row, col)
    eng.init()
    r, c = row, col
    for t in range(3):
        eng.apply(ins[t], dele[t])
        r, c = O.apply_updates(r, c, ins[t], dele[t])
    rel, nrm, l1 = pr_errors(eng.values(), O.pagerank(r, c))
    say(f"[gpu] deferred pr rel={rel:.2e} norm={nrm:.2e} pull_rounds={eng.stats()['pull_rounds']}")
    eng.close()
    assert rel <= PR_REL_TOL and nrm <= PR_NORM_TOL


def test_pr_graph_rounds_with_pull():
    # rounds submitted as HIP graphs (round kernel + pull + fix-up nodes per
    # round), replayed on every snapshot with host waits between them
    nv, row, col = sgraph.gen_rmat(13, 16, seed=61)
    ins, dele = sgraph.gen_updates(row, col, 13, 3, 2000, 2000, seed=62, hub_frac=0.8, hub_top=0.01)
    st = pr_stream(row, col, list(zip(ins, dele)), flags=sgraph.FLAG_GRAPH_ROUNDS)
    assert st["pull_rounds"] > 0


def test_c1_full_size():
    # C1: R-MAT scale 16 EF16, 2 snapshots of 10,000 fresh inserts + 10,000 deletes, SEGMENT 64, resident
    nv, row, col = sgraph.gen_rmat(16, 16, seed=sgraph.DEFAULT_SEED)
    ins, dele = sgraph.gen_updates(row, col, 16, 2, 10000, 10000, seed=sgraph.DEFAULT_SEED + 1)
    say(f"[gpu] C1 nv={nv} ne={len(col)}")
    pr_stream(row, col, list(zip(ins, dele)), segments=64)


def trav_modes(algo, scale, ef, nsnap, nupd, seed, **cfg):
    nv, row, col = sgraph.gen_rmat(scale, ef, seed=seed)
    ins, dele = sgraph.gen_updates(row, col, scale, nsnap, nupd, nupd, seed=seed + 1, hub_frac=0.5, hub_top=0.01)
    flags = sgraph.FLAG_IN_EDGES | cfg.pop("flags", 0)
    src = 0 if algo != "cc" else None
    graphs = [(row, col)]
    for t in range(nsnap):
        graphs.append(O.apply_updates(*graphs[-1], ins[t], dele[t]))
    res = {}
    for mode, div in (("never", sgraph.PULL_NEVER), ("always", sgraph.PULL_ALWAYS), ("default", 0)):
        e = sgraph.Engine(algo, weight_mode="seeded", weight_seed=7, flags=flags, pull_divisor=div, **cfg)
        try:
            e.load_csr(row, col)
            e.init(src)
            outs, pulls = [(e.values(), e.gather(sgraph.Q_PARENT))], e.stats()["pull_rounds"]
            for t in range(nsnap):
                e.apply(ins[t], dele[t])
                outs.append((e.values(), e.gather(sgraph.Q_PARENT)))
                pulls += e.stats()["pull_rounds"]
        finally:
            e.close()
        res[mode] = (outs, pulls)
        say(f"[gpu] {algo} pull mode {mode}: {pulls} pull rounds {cfg}")
    for t, (r, c) in enumerate(graphs):
        exp = O.bfs(r, c, 0) if algo == "bfs" else O.sssp(r, c, 0, "seeded", 7) if algo == "sssp" else O.cc(r, c)
        for mode in res:
            d, p = res[mode][0][t]
            assert np.array_equal(d, exp), f"G{t} {mode}: values differ from the oracle"
            if algo == "cc":
                assert O.cc_certificate(r, c, d, p) == 0, f"G{t} {mode}: CC certificate"
            else:
                # a unique fixed point: the smallest parent id of the best value
                assert np.array_equal(p, res["never"][0][t][1]), f"G{t} {mode}: parents differ from push rounds"
    assert res["never"][1] == 0 and res["always"][1] > 0
    return res


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_traversal_pull_rounds(algo):
    # direction-optimising rounds over the in-edge store: every round pulls,
    # none does, or the default share; same values (and SSSP/BFS parents)
    trav_modes(algo, 14, 16, 3, 3000, seed=67)


@pytest.mark.parametrize("algo", ["sssp", "cc"])
def test_traversal_pull_rounds_host_edges(algo):
    trav_modes(algo, 12, 16, 2, 2000, seed=69, flags=sgraph.FLAG_HOST_EDGES, segments=7, cache_level=1)


def test_traversal_pull_rounds_as_graphs():
    trav_modes("bfs", 13, 16, 2, 2000, seed=79, flags=sgraph.FLAG_GRAPH_ROUNDS)

