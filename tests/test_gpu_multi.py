"""Batched BFS/SSSP queries on one shared graph (SG_ALGO_MULTI): every query's
distances bit-exact against the oracle after the static solve and after each
snapshot, and distances and parents equal to a single-source engine's for
the same source (the same result alone and in the batch)."""
import numpy as np
import pytest

import oracle as O
import sgraph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def expected(kind, row, col, src, wmode, seed):
    if kind == "bfs":
        return O.bfs(row, col, src)
    return O.sssp(row, col, src, wmode, seed)


def check_all(eng, row, col, sources, kinds, wmode, seed, tag):
    for qi, (s, k) in enumerate(zip(sources, kinds)):
        got = eng.query(qi)
        exp = expected(k, row, col, int(s), wmode, seed)
        bad = np.nonzero(got != exp)[0]
        assert len(bad) == 0, f"{tag} q{qi} ({k} from {s}): {len(bad)} mismatches, first {bad[:5]} got {got[bad[:5]]} exp {exp[bad[:5]]}"


def run_multi(scale, ef, nsnap, nupd, sources, kinds, wmode="seeded", seed=9, gseed=41, flags=0):
    nv, row, col = sgraph.gen_rmat(scale, ef, seed=gseed)
    ins, dele = sgraph.gen_updates(row, col, scale, nsnap, nupd, nupd, seed=gseed + 1)
    say(f"[gpu] multi nv={nv} ne={len(col)} nq={len(sources)} snaps={nsnap} w={wmode} flags={flags}")
    e = sgraph.Engine("multi", weight_mode=wmode, weight_seed=seed, flags=flags)
    try:
        e.load_csr(row, col)
        e.init_queries(sources, kinds)
        check_all(e, row, col, sources, kinds, wmode, seed, "G0")
        r, c = row, col
        for t in range(nsnap):
            e.apply(ins[t], dele[t])
            r, c = O.apply_updates(r, c, ins[t], dele[t])
            check_all(e, r, c, sources, kinds, wmode, seed, f"G{t + 1}")
        return e.stats()
    finally:
        e.close()


@pytest.mark.parametrize("nq", [1, 2, 3, 12, 20, 64, 65])  # QS 1, 2, 4, 16, 32, 64, 64 x 2 groups
def test_query_counts(nq):
    rng = np.random.default_rng(nq)
    nv = 1 << 10
    sources = rng.integers(0, nv, nq).astype(np.uint32)
    sources[0] = 0
    kinds = ["sssp" if i % 3 else "bfs" for i in range(nq)]
    run_multi(10, 8, 3, 200, sources, kinds)


def test_graph_rounds_batch():
    rng = np.random.default_rng(9)
    sources = rng.integers(0, 1 << 11, 40).astype(np.uint32)
    kinds = ["sssp" if i % 2 else "bfs" for i in range(40)]
    run_multi(11, 8, 4, 400, sources, kinds, flags=sgraph.FLAG_GRAPH_ROUNDS)


def test_mixed_batch_hash128_with_duplicate_sources():
    sources = np.array([0, 5, 5, 17, 0, 1023, 300], dtype=np.uint32)
    kinds = ["sssp", "bfs", "sssp", "bfs", "bfs", "sssp", "sssp"]
    run_multi(10, 8, 4, 300, sources, kinds, wmode="hash128")


def test_larger_stream_128_queries():
    rng = np.random.default_rng(5)
    sources = rng.integers(0, 1 << 13, 128).astype(np.uint32)
    kinds = ["bfs"] * 64 + ["sssp"] * 64
    st = run_multi(13, 16, 3, 2000, sources, kinds)
    assert st["snapshots"] == 3


def test_same_result_alone_and_in_batch():
    nv, row, col = sgraph.gen_rmat(11, 8, seed=77)
    ins, dele = sgraph.gen_updates(row, col, 11, 3, 400, 400, seed=78)
    rng = np.random.default_rng(3)
    sources = rng.integers(0, nv, 40).astype(np.uint32)
    kinds = ["sssp" if i % 2 else "bfs" for i in range(40)]
    m = sgraph.Engine("multi", weight_mode="seeded", weight_seed=4)
    m.load_csr(row, col)
    m.init_queries(sources, kinds)
    singles = []
    for qi in (0, 1, 7, 38, 39):
        s = sgraph.Engine(kinds[qi], weight_mode="seeded", weight_seed=4)
        s.load_csr(row, col)
        s.init(int(sources[qi]))
        singles.append((qi, s))
    try:
        for t in range(3):
            m.apply(ins[t], dele[t])
            for qi, s in singles:
                s.apply(ins[t], dele[t])
                assert np.array_equal(m.query(qi), s.values()), f"G{t + 1} q{qi}: distances differ"
                assert np.array_equal(m.query(qi, sgraph.Q_PARENT), s.gather(sgraph.Q_PARENT)), f"G{t + 1} q{qi}: parents differ"
    finally:
        m.close()
        for _, s in singles:
            s.close()


def test_dropped_queries_leave_a_graph_only_engine():
    # nq = 0 drops the queries; later snapshots only update the graph, and new
    # queries on it match the oracle on the latest snapshot
    nv, row, col = sgraph.gen_rmat(11, 8, seed=61)
    ins, dele = sgraph.gen_updates(row, col, 11, 4, 400, 400, seed=62)
    m = sgraph.Engine("multi", weight_mode="seeded", weight_seed=3)
    try:
        m.load_csr(row, col)
        m.init_queries([], [])  # graph only from the start
        r, c = row, col
        m.apply(ins[0], dele[0])
        r, c = O.apply_updates(r, c, ins[0], dele[0])
        m.init_queries([0, 9], ["bfs", "sssp"])
        check_all(m, r, c, [0, 9], ["bfs", "sssp"], "seeded", 3, "G1")
        m.init_queries([], [])
        with pytest.raises(sgraph.SgError):
            m.query(0)
        for t in (1, 2, 3):
            m.apply(ins[t], dele[t])
            r, c = O.apply_updates(r, c, ins[t], dele[t])
            assert m.stats()["rounds"] == 0
        m.init_queries([5, 5, 100], ["sssp", "bfs", "bfs"])
        check_all(m, r, c, [5, 5, 100], ["sssp", "bfs", "bfs"], "seeded", 3, "G4")
    finally:
        m.close()


def test_errors():
    nv, row, col = sgraph.gen_rmat(8, 4, seed=1)
    s = sgraph.Engine("sssp")
    s.load_csr(row, col)
    with pytest.raises(sgraph.SgError):
        s.init_queries([0], ["sssp"])
    s.close()
    m = sgraph.Engine("multi")
    m.load_csr(row, col)
    with pytest.raises(sgraph.SgError):
        m.init_queries([nv], ["sssp"])
    with pytest.raises(sgraph.SgError):
        m.init_queries([0], ["cc"])
    m.init_queries([0, 1], ["bfs", "sssp"])
    with pytest.raises(sgraph.SgError):
        m.query(2)
    assert np.array_equal(m.query(0), O.bfs(row, col, 0))
    m.close()
