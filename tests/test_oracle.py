"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker:
golden fixtures, brute force, closed forms, and cross-checks between
independent restatements. CPU only."""
import json
import os

import numpy as np
import pytest

import oracle as O

from conftest import GOLDEN


def csr_from_adj(adj):
    row = np.zeros(len(adj) + 1, dtype=np.uint64)
    row[1:] = np.cumsum([len(a) for a in adj])
    col = np.array([v for a in adj for v in a], dtype=np.uint32)
    return row, col


def adj_from_csr(row, col):
    return [col[row[u]:row[u + 1]].tolist() for u in range(len(row) - 1)]


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


README = load("readme_expected.json")
TINY = load("tiny_expected.json")


def check_snapshot(row, col, exp, src, seed):
    assert O.bfs(row, col, src).tolist() == exp["bfs"]
    assert O.sssp(row, col, src, "hash128").tolist() == exp["sssp_hash128"]
    assert O.sssp(row, col, src, "seeded", seed).tolist() == exp["sssp_seeded"]
    assert O.sssp_fifo(row, col, src, "hash128").tolist() == exp["sssp_hash128"]
    assert O.cc(row, col).tolist() == exp["cc"]
    assert O.cc_label_prop(row, col).tolist() == exp["cc"]
    np.testing.assert_allclose(O.pagerank(row, col), exp["pr"], rtol=0, atol=1e-11)


def test_readme_fixture_all_snapshots():
    snaps = README["snapshots"]
    row, col = csr_from_adj(snaps[0]["adj"])
    check_snapshot(row, col, snaps[0], 0, README["seed"])
    ins = [(0, 3), (0, 4), (1, 4), (2, 3)]
    dele = [(0, 1), (0, 2), (1, 2), (1, 3), (2, 4), (3, 4)]
    for t in range(2):
        row, col = O.apply_updates(row, col, ins[2 * t:2 * t + 2], dele[3 * t:3 * t + 3])
        assert adj_from_csr(row, col) == snaps[t + 1]["adj"]
        check_snapshot(row, col, snaps[t + 1], 0, README["seed"])


def test_readme_values_match_survey():
    # G_1: BFS (0,inf,inf,1,1), SSSP d(3)=4 d(4)=5, labels (0,1,2,0,0)
    s1 = README["snapshots"][1]
    assert s1["bfs"] == [0, O.INF, O.INF, 1, 1]
    assert s1["sssp_hash128"][3:] == [4, 5]
    assert s1["cc"] == [0, 1, 2, 0, 0]


@pytest.mark.parametrize("case", range(len(TINY)))
def test_tiny_cases(case):
    c = TINY[case]
    snaps = c["snapshots"]
    row, col = csr_from_adj(snaps[0]["adj"])
    check_snapshot(row, col, snaps[0], c["source"], c["seed"])
    for t, up in enumerate(c["updates"]):
        row, col = O.apply_updates(row, col, up["ins"], up["del"])
        assert adj_from_csr(row, col) == snaps[t + 1]["adj"]
        check_snapshot(row, col, snaps[t + 1], c["source"], c["seed"])


def test_planted_weight_mistake_is_caught():
    # the CPU checker's weight without "+1" (hybrid_sssp_host.cu:52) must differ
    row, col = csr_from_adj(README["snapshots"][1]["adj"])
    assert O.sssp(row, col, 0, "hash128_noplus1").tolist() != README["snapshots"][1]["sssp_hash128"]


def test_update_order_inserts_then_deletes():
    # inserts are applied before deletes (framework.cuh:1244-1245): an edge
    # inserted and deleted in one snapshot is absent; the deleted-first order
    # would keep it, so this pins the order
    row, col = csr_from_adj([[1], [], []])
    r1, c1 = O.apply_updates(row, col, [(0, 2)], [(0, 2)])
    assert adj_from_csr(r1, c1) == [[1], [], []]
    # every copy of a deleted target goes; the order of the rest is kept
    row, col = csr_from_adj([[2, 1, 2, 3, 2], [], [], []])
    r2, c2 = O.apply_updates(row, col, [(0, 2), (0, 1)], [(0, 2)])
    assert adj_from_csr(r2, c2) == [[1, 3, 1], [], [], []]


def test_closed_forms():
    n = 50
    path = [[i + 1] if i + 1 < n else [] for i in range(n)]
    row, col = csr_from_adj(path)
    assert O.bfs(row, col, 0).tolist() == list(range(n))
    d = O.sssp(row, col, 0, "hash128").tolist()
    assert d == [sum((i + i + 1) % 128 + 1 for i in range(k)) for k in range(n)]
    assert O.cc(row, col).tolist() == [0] * n
    # cycle: every rank is 1 (r = 0.15 + 0.85 r)
    cyc = [[(i + 1) % n] for i in range(n)]
    row, col = csr_from_adj(cyc)
    np.testing.assert_allclose(O.pagerank(row, col), 1.0, atol=1e-12)
    # in-star: k leaves -> hub; r_leaf = 0.15, r_hub = 0.15 + 0.85 * k * 0.15
    k = 7
    star = [[k]] * k + [[]]
    row, col = csr_from_adj(star)
    r = O.pagerank(row, col)
    np.testing.assert_allclose(r[:k], 0.15, atol=1e-12)
    assert abs(r[k] - (0.15 + 0.85 * k * 0.15)) < 1e-12
    # reverse path labels: vertex i reached only from i..n-1 along i+1 -> i
    rev = [[i - 1] if i > 0 else [] for i in range(n)]
    row, col = csr_from_adj(rev)
    assert O.cc(row, col).tolist() == list(range(n))


def random_graph(rng, n, m):
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    order = np.lexsort((v[keep], u[keep]))
    u, v = u[keep][order], v[keep][order]
    row = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(row, u + 1, 1)
    return np.cumsum(row).astype(np.uint64), v.astype(np.uint32)


def test_floyd_warshall_and_closure_brute_force():
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = int(rng.integers(2, 30))
        row, col = random_graph(rng, n, int(rng.integers(0, 4 * n)))
        INF = float("inf")
        D = np.full((n, n), INF)
        for u in range(n):
            D[u, u] = 0
            for e in range(int(row[u]), int(row[u + 1])):
                D[u, col[e]] = min(D[u, col[e]], (u + int(col[e])) % 128 + 1)
        for k in range(n):
            D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
        src = int(rng.integers(0, n))
        exp = [int(x) if x < INF else O.INF for x in D[src]]
        assert O.sssp(row, col, src, "hash128").tolist() == exp
        reach = D < INF
        assert O.cc(row, col).tolist() == [int(np.argmax(reach[:, v])) for v in range(n)]


def test_pagerank_dense_solve():
    rng = np.random.default_rng(3)
    for trial in range(10):
        n = int(rng.integers(2, 40))
        row, col = random_graph(rng, n, int(rng.integers(0, 5 * n)))
        M = np.zeros((n, n))
        for u in range(n):
            d = int(row[u + 1] - row[u])
            for e in range(int(row[u]), int(row[u + 1])):
                M[col[e], u] += 1.0 / d
        exact = np.linalg.solve(np.eye(n) - 0.85 * M, np.full(n, 0.15))
        np.testing.assert_allclose(O.pagerank(row, col), exact, atol=1e-11)


def test_restatements_agree_on_rmat():
    import sgraph
    nv, row, col = sgraph.gen_rmat(12, 8)
    for mode in ("hash128", "seeded"):
        assert np.array_equal(O.sssp(row, col, 0, mode, 99), O.sssp_fifo(row, col, 0, mode, 99))
    assert np.array_equal(O.cc(row, col), O.cc_label_prop(row, col))
    lv = O.bfs(row, col, 0)
    d1 = O.sssp(row, col, 0, "one")
    assert np.array_equal(lv, d1)
    exact = O.pagerank(row, col)
    host = O.pagerank_host_f32(row, col)
    assert np.max(np.abs(host - exact) / exact) < 1e-3


def test_pma_geometry_hand_values():
    # compute_capacity (csr_graph.cuh:364-382)
    assert O.pma_geometry(4194304) == (32, 131072)
    assert O.pma_geometry(65536) == (16, 4096)
    assert O.pma_geometry(5) == (3, 2)
    assert O.pma_geometry(1) == (1, 1)


def test_segment_partition_hand_values():
    off = np.array([0, 4, 8, 12, 16], dtype=np.uint64)
    assert O.segment_partition(off, 2).tolist() == [0, 2, 4]
    off = np.array([0, 10, 11, 12, 13], dtype=np.uint64)
    # per = ceil(13/3) = 5: vertex 0 alone fills segment 0, then 1..3 (3 slots) end the walk
    assert O.segment_partition(off, 3).tolist() == [0, 1, 4]


# ---- certificates: pass on the oracles' own answers, fail on planted mistakes ----
def _rand_csr(nv, ne, seed):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, nv, ne)
    v = (rng.pareto(1.2, ne) * 3).astype(np.int64) % nv  # skewed targets, hubs included
    keep = u != v
    k = np.unique(u[keep].astype(np.int64) << 32 | v[keep])
    src, dst = (k >> 32).astype(np.int64), (k & 0xFFFFFFFF).astype(np.uint32)
    row = np.zeros(nv + 1, dtype=np.uint64)
    np.cumsum(np.bincount(src, minlength=nv), out=row[1:])
    return nv, row, dst


def _sp_parents(row, col, dist, wmode, seed):
    nv = len(row) - 1
    par = np.full(nv, O.INF, dtype=np.uint32)
    for u in range(nv):
        if dist[u] == O.INF:
            continue
        for e in range(int(row[u]), int(row[u + 1])):
            v = int(col[e])
            if par[v] == O.INF and v != 0 and dist[v] != O.INF and int(dist[u]) + O.weight(wmode, seed, u, v) == int(dist[v]):
                par[v] = u
    return par


def _cc_parents(row, col, label):
    # BFS from every root (label == own id) inside its label class
    nv = len(row) - 1
    par = np.full(nv, O.INF, dtype=np.uint32)
    seen = np.zeros(nv, dtype=bool)
    for r in range(nv):
        if label[r] != r:
            continue
        seen[r] = True
        q = [r]
        while q:
            u = q.pop(0)
            for e in range(int(row[u]), int(row[u + 1])):
                v = int(col[e])
                if not seen[v] and label[v] == r:
                    seen[v] = True
                    par[v] = u
                    q.append(v)
    return par


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sp_certificate_accepts_distances_and_rejects_mistakes(seed):
    nv, row, col = _rand_csr(256, 1500, seed)
    for wmode in ("seeded", "one"):
        d = O.sssp(row, col, 0, wmode, 11) if wmode != "one" else O.bfs(row, col, 0)
        p = _sp_parents(row, col, d, wmode, 11)
        assert O.sp_certificate(row, col, 0, d, p, wmode, 11) == 0
        reached = np.nonzero((d != O.INF) & (np.arange(nv) != 0))[0]
        v = int(reached[len(reached) // 2])
        for delta in (-1, +1):
            bad = d.copy()
            bad[v] = int(bad[v]) + delta
            assert O.sp_certificate(row, col, 0, bad, p, wmode, 11) > 0
        badp = p.copy()
        badp[v] = (int(p[v]) + 1) % nv
        assert O.sp_certificate(row, col, 0, d, badp, wmode, 11) > 0


@pytest.mark.parametrize("seed", [4, 5])
def test_cc_certificate_accepts_labels_and_rejects_mistakes(seed):
    nv, row, col = _rand_csr(256, 500, seed)
    lab = O.cc(row, col)
    par = _cc_parents(row, col, lab)
    assert O.cc_certificate(row, col, lab, par) == 0
    movers = np.nonzero(lab != np.arange(nv))[0]
    v = int(movers[0])
    low = lab.copy()
    low[v] = 0 if lab[v] != 0 else 1  # below its min ancestor (or a non-ancestor)
    assert O.cc_certificate(row, col, low, par) > 0
    high = lab.copy()
    high[v] = v  # its own id: an in-edge from a smaller label now violates
    assert O.cc_certificate(row, col, high, par) > 0
    cyc = par.copy()
    cyc[int(par[v])] = v  # a parent cycle
    assert O.cc_certificate(row, col, lab, cyc) > 0


def test_pr_residual_bounds_the_error():
    nv, row, col = _rand_csr(512, 4000, 6)
    exact = O.pagerank(row, col)
    l1, mx, _ = O.pr_residual(row, col, exact.astype(np.float32))
    assert mx < 1e-5 and l1 < 1e-3
    rng = np.random.default_rng(7)
    for _ in range(5):
        r = (exact * (1 + 0.05 * rng.standard_normal(nv))).astype(np.float32)
        l1, mx, _ = O.pr_residual(row, col, r)
        err = float(np.abs(r.astype(np.float64) - exact).sum())
        assert err <= l1 / (1 - 0.85) * (1 + 1e-9)


# ---- widest path (the operator library's program; no reference counterpart) ----
def _widest_brute(nv, row, col, src, wmode, seed):
    """Largest bottleneck over every simple path (depth-first enumeration)."""
    best = [0] * nv
    best[src] = O.INF
    w = {}
    for u in range(nv):
        for e in range(int(row[u]), int(row[u + 1])):
            v = int(col[e])
            w[(u, v)] = max(w.get((u, v), 0), O.weight(wmode, seed, u, v))

    def walk(u, width, on_path):
        for (a, b), ww in w.items():
            if a != u or b in on_path:
                continue
            nw = min(width, ww)
            best[b] = max(best[b], nw)
            on_path.add(b)
            walk(b, nw, on_path)
            on_path.discard(b)

    walk(src, O.INF, {src})
    return np.array(best, dtype=np.uint32)


def _widest_fixed_point(row, col, src, wmode, seed):
    """width[v] = max(width[v], max over edges u -> v of min(width[u], w)) until stable."""
    nv = len(row) - 1
    s, w = O.edge_weights(row, col, wmode, seed)
    width = np.zeros(nv, dtype=np.uint64)
    width[src] = O.INF
    while True:
        offer = np.minimum(width[s], w.astype(np.uint64))
        nw = width.copy()
        np.maximum.at(nw, col, offer)
        if np.array_equal(nw, width):
            return width.astype(np.uint32)
        width = nw


def _widest_parents(row, col, src, width, wmode, seed):
    """A parent tree for the widths from a pure-Python widest-first search
    (heapq): a vertex's parent is the popped vertex whose edge last widened it.
    Its widths must equal the given ones."""
    import heapq
    nv = len(row) - 1
    best = [0] * nv
    par = np.full(nv, O.INF, dtype=np.uint32)
    best[src] = O.INF
    heap, done = [(-O.INF, src)], [False] * nv
    while heap:
        nw, u = heapq.heappop(heap)
        if done[u]:
            continue
        done[u] = True
        for e in range(int(row[u]), int(row[u + 1])):
            v = int(col[e])
            c = min(best[u], O.weight(wmode, seed, u, v))
            if c > best[v] and not done[v]:
                best[v], par[v] = c, u
                heapq.heappush(heap, (-c, v))
    assert np.array_equal(np.array(best, dtype=np.uint32), width)
    return par


@pytest.mark.parametrize("seed", range(6))
def test_widest_path_brute_force_tiny(seed):
    rng = np.random.default_rng(300 + seed)
    nv = int(rng.integers(2, 8))
    ne = int(rng.integers(0, 3 * nv))
    src = rng.integers(0, nv, ne).astype(np.uint32)
    dst = rng.integers(0, nv, ne).astype(np.uint32)  # self-loops and parallel edges included
    order = np.lexsort((dst, src))
    row = np.zeros(nv + 1, dtype=np.uint64)
    np.cumsum(np.bincount(src, minlength=nv), out=row[1:])
    col = dst[order]
    for wmode in ("hash128", "seeded", "one"):
        for s0 in range(nv):
            exp = _widest_brute(nv, row, col, s0, wmode, 5)
            assert np.array_equal(O.widest_path(row, col, s0, wmode, 5), exp), (seed, wmode, s0)


def test_widest_path_fixed_point_and_certificate():
    rng = np.random.default_rng(7)  # uniform random digraph, mean degree 6: most vertices reached
    nv, ne = 512, 3072
    k = np.unique(rng.integers(0, nv, ne).astype(np.int64) << 32 | rng.integers(0, nv, ne))
    row = np.zeros(nv + 1, dtype=np.uint64)
    np.cumsum(np.bincount(k >> 32, minlength=nv), out=row[1:])
    col = (k & 0xFFFFFFFF).astype(np.uint32)
    src = 3
    for wmode in ("seeded", "hash128"):
        wd = O.widest_path(row, col, src, wmode, 13)
        assert np.array_equal(wd, _widest_fixed_point(row, col, src, wmode, 13))
        assert (wd > 0).sum() > 100
        par = _widest_parents(row, col, src, wd, wmode, 13)
        assert O.widest_certificate(row, col, src, wd, par, wmode, 13) == 0
        # a reached vertex two parent steps below the source
        deep = [int(v) for v in np.nonzero((wd > 0) & (par != O.INF))[0] if par[v] != src and par[par[v]] != O.INF]
        v = deep[len(deep) // 2]
        for delta in (-1, +1):  # a narrower or wider width
            bad = wd.copy()
            bad[v] = int(bad[v]) + delta
            assert O.widest_certificate(row, col, src, bad, par, wmode, 13) > 0
        cyc = par.copy()
        cyc[par[v]] = v  # a 2-cycle of parents
        assert O.widest_certificate(row, col, src, wd, cyc, wmode, 13) > 0
        badp = par.copy()
        badp[v] = O.INF  # a reached vertex without a parent
        assert O.widest_certificate(row, col, src, wd, badp, wmode, 13) > 0
