"""Parity of the HIP engine (through the C ABI) with the CPU oracle, per
snapshot: BFS levels, SSSP distances and min-ancestor labels bit-exact;
PageRank within a relative error of 5e-5 of the FP64 fixed point and 1e-6
max-abs on L1-normalised ranks; the PMA adjacency sequences equal to the
update model's (inserts appended in file order, then every deleted target
removed) after every snapshot."""
import json
import os

import numpy as np
import pytest

import oracle as O
import sgraph
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

PR_REL_TOL = 5e-5     # max |r - r*| / r* on raw ranks
PR_NORM_TOL = 1e-6    # max |r/sum(r) - r*/sum(r*)|


def say(*a):
    print(*a, flush=True)


def expected(algo, row, col, src, wmode, seed):
    if algo == "bfs":
        return O.bfs(row, col, src)
    if algo == "sssp":
        return O.sssp(row, col, src, wmode, seed)
    if algo == "cc":
        return O.cc(row, col)
    return O.pagerank(row, col)


def check_values(algo, eng, row, col, src, wmode, seed, tag):
    got = eng.values()
    exp = expected(algo, row, col, src, wmode, seed)
    if algo == "pr":
        rel = np.max(np.abs(got.astype(np.float64) - exp) / exp) if len(exp) else 0.0
        gn = got.astype(np.float64) / got.astype(np.float64).sum()
        en = exp / exp.sum()
        assert rel <= PR_REL_TOL, f"{tag}: PR rel err {rel}"
        assert np.max(np.abs(gn - en)) <= PR_NORM_TOL, f"{tag}: PR normalised err"
    else:
        bad = np.nonzero(got != exp)[0]
        assert len(bad) == 0, f"{tag}: {len(bad)} mismatches, first {bad[:5]} got {got[bad[:5]]} exp {exp[bad[:5]]}"


def check_pma(eng, row, col, tag):
    r, c, off, deg, slots = eng.adjacency()
    assert np.array_equal(r, row), f"{tag}: degrees differ"
    assert np.array_equal(c, col), f"{tag}: adjacency sequences differ"
    # invariants: ordered ranges, live prefix, gaps elsewhere
    cap = len(slots)
    assert off[-1] == cap and np.all(np.diff(off.astype(np.int64)) >= 0)
    assert np.all(off[:-1] + deg <= off[1:])
    live = np.zeros(cap, dtype=bool)
    d = deg.astype(np.int64)
    idx = np.repeat(off[:-1].astype(np.int64), d) + (np.arange(int(d.sum())) - np.repeat(np.cumsum(d) - d, d))
    live[idx] = True
    assert np.all(slots[~live] == 0xFFFFFFFF), f"{tag}: non-gap value outside live ranges"


def run_stream(algo, row, col, snaps, src=0, wmode="hash128", seed=0, pma=True, **cfg):
    say(f"[gpu] {algo} nv={len(row) - 1} ne={len(col)} snaps={len(snaps)} w={wmode} cfg={cfg}")
    eng = sgraph.Engine(algo, weight_mode=wmode, weight_seed=seed, **cfg)
    try:
        eng.load_csr(row, col)
        if pma:
            check_pma(eng, row, col, "load")
        eng.init(src if algo in ("bfs", "sssp") else None)
        check_values(algo, eng, row, col, src, wmode, seed, f"{algo} G0")
        r, c = row, col
        for t, (ins, dele) in enumerate(snaps):
            eng.apply(ins, dele)
            r, c = O.apply_updates(r, c, ins, dele)
            if pma:
                check_pma(eng, r, c, f"{algo} G{t + 1}")
            check_values(algo, eng, r, c, src, wmode, seed, f"{algo} G{t + 1}")
        return eng.stats()
    finally:
        eng.close()


ALGOS = ["bfs", "sssp", "cc", "pr"]


@pytest.mark.parametrize("algo", ALGOS)
def test_readme_fixture(algo):
    nv, row, col = sgraph.read_edge_list(os.path.join(GOLDEN, "readme_graph.txt"))
    ins, dele = sgraph.read_updates(os.path.join(GOLDEN, "readme_update.txt"))
    sizes = sgraph.read_stream_sizes(os.path.join(GOLDEN, "readme_stream_size.txt"))
    snaps, io, do = [], 0, 0
    for a, d in sizes.tolist():
        snaps.append((ins[io:io + a], dele[do:do + d]))
        io += a
        do += d
    run_stream(algo, row, col, snaps)
    exp = json.load(open(os.path.join(GOLDEN, "readme_expected.json")))["snapshots"][2]
    eng = sgraph.Engine(algo)
    eng.load_csr(row, col)
    eng.init(0 if algo in ("bfs", "sssp") else None)
    for s in snaps:
        eng.apply(*s)
    key = {"bfs": "bfs", "sssp": "sssp_hash128", "cc": "cc", "pr": "pr"}[algo]
    if algo == "pr":
        np.testing.assert_allclose(eng.values(), exp[key], rtol=PR_REL_TOL)
    else:
        assert eng.values().tolist() == exp[key]
    eng.close()


@pytest.mark.parametrize("algo", ALGOS)
def test_tiny_cases(algo):
    cases = json.load(open(os.path.join(GOLDEN, "tiny_expected.json")))
    for c in cases:
        adj = c["snapshots"][0]["adj"]
        row = np.zeros(c["nv"] + 1, dtype=np.uint64)
        row[1:] = np.cumsum([len(a) for a in adj])
        col = np.array([v for a in adj for v in a], dtype=np.uint32)
        snaps = [(np.array(u["ins"], dtype=np.uint32).reshape(-1, 2), np.array(u["del"], dtype=np.uint32).reshape(-1, 2))
                 for u in c["updates"]]
        run_stream(algo, row, col, snaps, src=c["source"], wmode="seeded" if algo == "sssp" else "hash128",
                   seed=c["seed"])


@pytest.mark.parametrize("algo", ALGOS)
def test_rmat_stream(algo):
    nv, row, col = sgraph.gen_rmat(12, 8, seed=11)
    ins, dele = sgraph.gen_updates(row, col, 12, 4, 600, 600, seed=12)
    run_stream(algo, row, col, list(zip(ins, dele)), wmode="seeded" if algo == "sssp" else "hash128", seed=99)


@pytest.mark.parametrize("algo", ALGOS)
def test_rmat_stream_graph_rounds(algo):
    # the same stream with every batch of rounds submitted as a HIP graph: the
    # graphs are replayed many times, with eager kernels and host waits between
    nv, row, col = sgraph.gen_rmat(12, 8, seed=11)
    ins, dele = sgraph.gen_updates(row, col, 12, 4, 600, 600, seed=12)
    run_stream(algo, row, col, list(zip(ins, dele)), wmode="seeded" if algo == "sssp" else "hash128", seed=99,
               flags=sgraph.FLAG_GRAPH_ROUNDS)


@pytest.mark.parametrize("timing", ["all", "sweep"])
def test_kernel_timing_classes(timing):
    # event timing changes nothing computed; SG_FLAG_TIME_SWEEP times the sweep alone
    nv, row, col = sgraph.gen_rmat(11, 8, seed=21)
    ins, dele = sgraph.gen_updates(row, col, 11, 3, 400, 400, seed=22)
    flags = sgraph.FLAG_KERNEL_TIMING | (sgraph.FLAG_TIME_SWEEP if timing == "sweep" else 0)
    e = sgraph.Engine("sssp", weight_mode="seeded", weight_seed=5, flags=flags)
    try:
        e.load_csr(row, col)
        e.init(0)
        r, c = row, col
        for t in range(3):
            e.apply(ins[t], dele[t])
            r, c = O.apply_updates(r, c, ins[t], dele[t])
            check_values("sssp", e, r, c, 0, "seeded", 5, f"timing={timing} G{t + 1}")
        kt = e.kernel_times()
        assert 1 <= kt["sweep"][0] <= 3 and kt["sweep"][1] > 0 and kt["sweep"][2] > 0  # one per snapshot that invalidates
        if timing == "sweep":
            assert kt["relax"][0] == 0 and kt["merge"][0] == 0
        else:
            assert kt["relax"][0] > 0 and kt["merge"][0] > 0
    finally:
        e.close()


@pytest.mark.parametrize("algo", ALGOS)
def test_checkpoint_restore_continues_the_stream(tmp_path, algo):
    # a restored engine holds the saved graph and state and gives the saved
    # engine's results on the later snapshots
    nv, row, col = sgraph.gen_rmat(11, 8, seed=71)
    ins, dele = sgraph.gen_updates(row, col, 11, 4, 300, 300, seed=72)
    kw = dict(weight_mode="seeded", weight_seed=5)
    src = 0 if algo in ("bfs", "sssp") else None
    a = sgraph.Engine(algo, **kw)
    b = sgraph.Engine(algo, **kw)
    try:
        a.load_csr(row, col)
        a.init(src)
        r, c = row, col
        for t in range(2):
            a.apply(ins[t], dele[t])
            r, c = O.apply_updates(r, c, ins[t], dele[t])
        path = str(tmp_path / "state.bin")
        a.save_state(path)
        b.restore_state(path)
        assert b.nv == nv and np.array_equal(b.values(), a.values())
        check_pma(b, r, c, f"{algo} restored")
        for t in (2, 3):
            a.apply(ins[t], dele[t])
            b.apply(ins[t], dele[t])
            r, c = O.apply_updates(r, c, ins[t], dele[t])
            check_values(algo, b, r, c, 0, "seeded", 5, f"{algo} restored G{t + 1}")
            if algo != "pr":
                assert np.array_equal(b.values(), a.values())
                if algo != "cc":
                    assert np.array_equal(b.gather(sgraph.Q_PARENT), a.gather(sgraph.Q_PARENT))
    finally:
        a.close()
        b.close()
    # the checkpoint only loads into a fresh engine of the same program (and,
    # for SSSP, the same weights)
    wrong = sgraph.Engine("bfs" if algo != "bfs" else "sssp", **kw)
    with pytest.raises(sgraph.SgError):
        wrong.restore_state(path)
    wrong.close()
    if algo == "sssp":
        w2 = sgraph.Engine("sssp", weight_mode="seeded", weight_seed=6)
        with pytest.raises(sgraph.SgError):
            w2.restore_state(path)
        w2.close()
    used = sgraph.Engine(algo, **kw)
    used.load_csr(row, col)
    with pytest.raises(sgraph.SgError):
        used.restore_state(path)
    used.close()


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_rmat_stream_larger(algo):
    nv, row, col = sgraph.gen_rmat(15, 16, seed=21)
    ins, dele = sgraph.gen_updates(row, col, 15, 3, 5000, 5000, seed=22)
    run_stream(algo, row, col, list(zip(ins, dele)), wmode="hash128", pma=(algo == "cc"))


def test_pr_permuted_larger():
    nv, row, col = sgraph.gen_rmat(14, 16, seed=31, permute=True)
    ins, dele = sgraph.gen_updates(row, col, 14, 3, 3000, 3000, seed=32)
    run_stream("pr", row, col, list(zip(ins, dele)), pma=False)


@pytest.mark.parametrize("algo", ["sssp", "pr"])
def test_dense_pma_forces_rebalance_and_resize(algo):
    # little slack and inserts concentrated on hubs: windows re-spread, capacity doubles
    nv, row, col = sgraph.gen_rmat(11, 8, seed=41)
    ins, dele = sgraph.gen_updates(row, col, 11, 5, 3000, 200, seed=42, hub_frac=0.9, hub_top=0.01)
    st = run_stream(algo, row, col, list(zip(ins, dele)), pma_density=0.9)
    say(f"[gpu] rebalances={st['rebalances']} resizes={st['resizes']}")
    assert st["rebalances"] > 0 and st["resizes"] > 0


@pytest.mark.parametrize("algo", ["bfs", "cc"])
def test_compaction_in_place_with_rebalance(algo):
    # many deletes next to hub-biased inserts in a dense PMA: short lists with
    # deletes are compacted where they lie during the batch's preparation, long
    # or overflowing ones are staged, and window re-spreads move both kinds
    nv, row, col = sgraph.gen_rmat(11, 8, seed=43)
    ins, dele = sgraph.gen_updates(row, col, 11, 4, 2500, 2000, seed=44, hub_frac=0.8, hub_top=0.02)
    st = run_stream(algo, row, col, list(zip(ins, dele)), pma_density=0.85)
    say(f"[gpu] rebalances={st['rebalances']} resizes={st['resizes']}")
    assert st["rebalances"] > 0


@pytest.mark.parametrize("algo", ALGOS)
def test_edge_cases(algo):
    # isolated source, a deleted-then-reinserted edge, an edge inserted and
    # deleted in one snapshot, duplicate inserts, deletes of absent edges,
    # deleting every out-edge of the source, empty snapshots
    nv = 9
    adj = [[1, 2], [3], [3, 4], [5], [5], [], [7], [], []]
    row = np.zeros(nv + 1, dtype=np.uint64)
    row[1:] = np.cumsum([len(a) for a in adj])
    col = np.array([v for a in adj for v in a], dtype=np.uint32)
    E = lambda *p: np.array(p, dtype=np.uint32).reshape(-1, 2)
    snaps = [
        (E(), E()),
        (E((0, 6), (0, 6), (4, 8)), E((0, 1), (8, 0), (3, 5))),
        (E((0, 1), (2, 8)), E((2, 8), (6, 7), (0, 6))),
        (E(), E((0, 1), (0, 2))),
        (E((8, 0), (0, 3)), E()),
    ]
    run_stream(algo, row, col, snaps, src=0)
    run_stream(algo, row, col, snaps, src=8)


def test_device_batches_equal_host_batches():
    nv, row, col = sgraph.gen_rmat(11, 8, seed=51)
    ins, dele = sgraph.gen_updates(row, col, 11, 3, 500, 500, seed=52)
    a = sgraph.Engine("sssp", weight_mode="seeded", weight_seed=3)
    b = sgraph.Engine("sssp", weight_mode="seeded", weight_seed=3)
    for e in (a, b):
        e.load_csr(row, col)
        e.init(0)
    for t in range(3):
        a.apply(ins[t], dele[t])
        di, dd = b.to_device(ins[t]), b.to_device(dele[t])
        b.apply_device(di, len(ins[t]), dd, len(dele[t]))
        b.device_free(di)
        b.device_free(dd)
        assert np.array_equal(a.values(), b.values())
        assert np.array_equal(a.gather(sgraph.Q_PARENT), b.gather(sgraph.Q_PARENT))
    a.close()
    b.close()


def test_invalid_ids():
    nv, row, col = sgraph.gen_rmat(8, 4, seed=61)
    e = sgraph.Engine("bfs")
    e.load_csr(row, col)
    e.init(0)
    before = e.values()
    with pytest.raises(sgraph.SgError) as x:
        e.apply(np.array([[0, nv]], dtype=np.uint32), np.zeros((0, 2), dtype=np.uint32))
    assert x.value.status == sgraph.SG_EINVAL
    assert np.array_equal(e.values(), before)
    # device path: out-of-range pairs are dropped and counted
    d = e.to_device(np.array([[0, nv], [nv + 5, 1]], dtype=np.uint32))
    e.apply_device(d, 2, 0, 0)
    assert e.stats()["dropped_updates"] == 2
    assert np.array_equal(e.values(), before)
    with pytest.raises(sgraph.SgError):
        e.init(nv)
    e.close()


def test_parents_form_shortest_path_tree():
    nv, row, col = sgraph.gen_rmat(12, 8, seed=71)
    ins, dele = sgraph.gen_updates(row, col, 12, 2, 800, 800, seed=72)
    e = sgraph.Engine("sssp", weight_mode="seeded", weight_seed=5)
    e.load_csr(row, col)
    e.init(0)
    r, c = row, col
    for t in range(2):
        e.apply(ins[t], dele[t])
        r, c = O.apply_updates(r, c, ins[t], dele[t])
    d = e.values()
    p = e.gather(sgraph.Q_PARENT)
    edges = set(zip(np.repeat(np.arange(nv), np.diff(r).astype(np.int64)).tolist(), c.tolist()))
    for v in range(nv):
        if v == 0 or d[v] == sgraph.INF:
            assert p[v] == sgraph.INF
            continue
        u = int(p[v])
        assert (u, v) in edges
        assert d[v] == d[u] + sgraph.edge_weight("seeded", 5, u, v)
    e.close()
