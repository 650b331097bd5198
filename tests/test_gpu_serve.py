"""Request loop (sgraph_serve.h): update batches, BFS/SSSP queries, CC label
lookups and PageRank top-k submitted from several threads while the worker
runs. Every answer must be the one for the snapshot it reports: BFS/SSSP
equal to the oracle and to a single-source engine on that snapshot (the same
result alone and in the mix), CC labels equal to the oracle's, top-k ranks
within the PageRank tolerance of the FP64 fixed point."""
import threading

import numpy as np
import pytest

import oracle as O
import sgraph

pytestmark = pytest.mark.gpu


def snapshots(row, col, ins, dele):
    graphs = [(row, col)]
    r, c = row, col
    for t in range(len(ins)):
        r, c = O.apply_updates(r, c, ins[t], dele[t])
        graphs.append((r, c))
    return graphs


def test_mixed_requests_match_their_snapshots():
    scale, nsnap = 10, 4
    nv, row, col = sgraph.gen_rmat(scale, 8, seed=91)
    ins, dele = sgraph.gen_updates(row, col, scale, nsnap, 150, 150, seed=92)
    graphs = snapshots(row, col, ins, dele)
    srv = sgraph.Server(row, col, query_slots=16, weight_mode="seeded", weight_seed=8)
    rng = np.random.default_rng(1)
    results = []
    lock = threading.Lock()

    def client(cid):
        r = np.random.default_rng(100 + cid)
        mine = []
        for i in range(12):
            kind = [sgraph.REQ_BFS, sgraph.REQ_SSSP, sgraph.REQ_CC, sgraph.REQ_PR_TOPK][int(r.integers(0, 4))]
            arg = int(r.integers(0, nv)) if kind != sgraph.REQ_PR_TOPK else 5
            mine.append((kind, arg, srv.submit(kind, arg)))
        for kind, arg, t in mine:
            out, info = srv.wait(t, kind, arg)
            with lock:
                results.append((kind, arg, out, info))

    threads = [threading.Thread(target=client, args=(i,)) for i in range(3)]
    for th in threads:
        th.start()
    # updates from one thread, in order: snapshot t is graphs[t]
    for t in range(nsnap):
        _, info = srv.wait(srv.submit(sgraph.REQ_UPDATE, ins=ins[t], dele=dele[t]))
        assert info["snapshot"] == t + 1
    for th in threads:
        th.join()
    st = srv.stats()
    srv.close()
    assert st["snapshots"] == nsnap and st["queries"] + st["lookups"] == len(results) == 36
    assert st["p99_ms"] >= st["p50_ms"] > 0
    singles = {}
    for kind, arg, out, info in results:
        r, c = graphs[info["snapshot"]]
        tag = f"kind {kind} arg {arg} snapshot {info['snapshot']}"
        if kind == sgraph.REQ_BFS:
            assert np.array_equal(out, O.bfs(r, c, arg)), tag
        elif kind == sgraph.REQ_SSSP:
            exp = O.sssp(r, c, arg, "seeded", 8)
            assert np.array_equal(out, exp), tag
            if info["snapshot"] not in singles:  # alone: a single-source engine on the same snapshot
                e = sgraph.Engine("sssp", weight_mode="seeded", weight_seed=8)
                e.load_csr(r, c)
                e.init(arg)
                assert np.array_equal(e.values(), out), tag
                e.close()
                singles[info["snapshot"]] = True
        elif kind == sgraph.REQ_CC:
            assert out[0] == O.cc(r, c)[arg], tag
        else:
            ids, ranks = out
            exp = O.pagerank(r, c)
            assert np.all(np.abs(ranks - exp[ids]) <= 5e-5 * exp[ids]), tag
            kth = np.sort(exp)[::-1][len(ids) - 1]
            assert ranks.min() >= kth * (1 - 1e-4), tag


def test_updates_before_the_first_query():
    nv, row, col = sgraph.gen_rmat(10, 8, seed=31)
    ins, dele = sgraph.gen_updates(row, col, 10, 2, 200, 200, seed=32)
    srv = sgraph.Server(row, col, query_slots=8)
    try:
        for t in range(2):  # no query has been seen: the batched engine holds only the graph
            _, info = srv.wait(srv.submit(sgraph.REQ_UPDATE, ins=ins[t], dele=dele[t]))
            assert info["snapshot"] == t + 1
        r, c = snapshots(row, col, ins, dele)[2]
        out, info = srv.wait(srv.submit(sgraph.REQ_BFS, 3), sgraph.REQ_BFS, 3)
        assert info["snapshot"] == 2 and np.array_equal(out, O.bfs(r, c, 3))
        out, info = srv.wait(srv.submit(sgraph.REQ_CC, 7), sgraph.REQ_CC, 7)
        assert out[0] == O.cc(r, c)[7]
    finally:
        srv.close()


def test_queries_are_batched_and_target_answers():
    nv, row, col = sgraph.gen_rmat(10, 8, seed=5)
    srv = sgraph.Server(row, col, query_slots=64, standing=0)
    tickets = [(s, srv.submit(sgraph.REQ_BFS, s, target=7)) for s in range(40)]
    for s, t in tickets:
        out, info = srv.wait(t, sgraph.REQ_BFS, s, target=7)
        assert out[0] == O.bfs(row, col, s)[7]
    st = srv.stats()
    assert st["max_batch"] > 1
    with pytest.raises(sgraph.SgError):
        srv.submit(sgraph.REQ_CC, 0)  # CC not maintained
    with pytest.raises(sgraph.SgError):
        srv.submit(sgraph.REQ_BFS, nv)
    srv.close()


def test_pointwise_gather_and_topk_on_engines():
    nv, row, col = sgraph.gen_rmat(10, 8, seed=13)
    e = sgraph.Engine("cc")
    e.load_csr(row, col)
    e.init()
    vs = np.array([0, 5, nv - 1, nv + 3], dtype=np.uint32)
    got = e.gather_at(vs)
    lab = O.cc(row, col)
    assert got[:3].tolist() == lab[vs[:3]].tolist() and got[3] == sgraph.INF
    e.close()
    p = sgraph.Engine("pr")
    p.load_csr(row, col)
    p.init()
    ids, ranks = p.topk(10)
    vals = p.values()
    assert np.array_equal(ranks, vals[ids])
    assert ranks.min() >= np.sort(vals)[::-1][9]
    p.close()


def test_deferred_pagerank_converges_once_for_many_snapshots():
    nv, row, col = sgraph.gen_rmat(11, 8, seed=17)
    ins, dele = sgraph.gen_updates(row, col, 11, 3, 300, 300, seed=18)
    lazy = sgraph.Engine("pr", flags=sgraph.FLAG_DEFER_CONVERGE)
    eager = sgraph.Engine("pr")
    for e in (lazy, eager):
        e.load_csr(row, col)
        e.init()
    r, c = row, col
    for t in range(3):
        lazy.apply(ins[t], dele[t])
        eager.apply(ins[t], dele[t])
        assert lazy.stats()["rounds"] == 0  # no rounds inside a deferred snapshot
        r, c = O.apply_updates(r, c, ins[t], dele[t])
    got = lazy.values()  # converges the three deferred snapshots
    exp = O.pagerank(r, c)
    assert np.all(np.abs(got - exp) <= 5e-5 * exp)
    assert np.all(np.abs(got - eager.values()) <= 5e-5 * exp)
    lazy.converge()  # nothing pending: a no-op
    assert np.array_equal(lazy.values(), got)
    lazy.close()
    eager.close()


def test_query_slots_join_mid_convergence():
    # continuous batching on the engine: queries join free slots while others
    # converge, leave when done, slots are reused; every answer equals the
    # oracle and a single-source engine (distances and parents)
    nv, row, col = sgraph.gen_rmat(14, 16, seed=97)
    ins, dele = sgraph.gen_updates(row, col, 14, 2, 2000, 2000, seed=98)
    e = sgraph.Engine("multi", weight_mode="seeded", weight_seed=4)
    e.load_csr(row, col)
    e.query_open(8)
    rng = np.random.default_rng(99)
    src = {}
    hubs = [int(x) for x in np.argsort(np.diff(row.astype(np.int64)))[-2:]]  # deep enough for one round not to end them
    first = e.query_add(hubs, ["sssp", "bfs"])
    src.update({int(s): (q, k) for s, q, k in zip(first, hubs, ["sssp", "bfs"])})
    done, active = e.query_step(1)  # one round: the first two are mid-convergence
    assert active == 2 and len(done) == 0
    more = [int(x) for x in rng.integers(0, nv, 4)]
    kinds = ["sssp", "bfs", "sssp", "bfs"]
    slots = e.query_add(more, kinds)
    src.update({int(s): (q, k) for s, q, k in zip(slots, more, kinds)})
    with pytest.raises(sgraph.SgError):  # an update waits for the running queries
        e.apply(ins[0], dele[0])
    results = {}
    while True:
        done, active = e.query_step(2)
        for s in done:
            results[int(s)] = (e.query(int(s)), e.query(int(s), sgraph.Q_PARENT))
        if len(done):
            e.query_release(done)
        if active == 0:
            break
    assert sorted(results) == sorted(src)
    for s, (d, p) in results.items():
        q, k = src[s]
        exp = O.bfs(row, col, q) if k == "bfs" else O.sssp(row, col, q, "seeded", 4)
        assert np.array_equal(d, exp), f"slot {s} ({k} from {q})"
        one = sgraph.Engine(k, weight_mode="seeded", weight_seed=4)
        one.load_csr(row, col)
        one.init(q)
        assert np.array_equal(one.values(), d) and np.array_equal(one.gather(sgraph.Q_PARENT), p)
        one.close()
    # no slot in use: an update only changes the graph; freed slots are reused
    e.apply(ins[0], dele[0])
    r1, c1 = O.apply_updates(row, col, ins[0], dele[0])
    s2 = e.query_add([3], ["sssp"])
    assert int(s2[0]) < 8
    finished = []
    while True:
        done, active = e.query_step(4)
        finished += [int(x) for x in done]
        if active == 0:
            break
    assert finished == [int(s2[0])]
    assert np.array_equal(e.query(int(s2[0])), O.sssp(r1, c1, 3, "seeded", 4))
    # a converged query still held is brought to the next snapshot by the update
    e.apply(ins[1], dele[1])
    r2, c2 = O.apply_updates(r1, c1, ins[1], dele[1])
    assert np.array_equal(e.query(int(s2[0])), O.sssp(r2, c2, 3, "seeded", 4))
    e.close()


def test_server_refuses_queries_without_hbm_for_them():
    # an HBM reserve larger than the device: no query slot can be opened, so
    # BFS/SSSP requests fail with SG_EBUDGET (counted), lookups still work
    nv, row, col = sgraph.gen_rmat(10, 8, seed=101)
    srv = sgraph.Server(row, col, query_slots=16, hbm_reserve_gb=1e6)
    with pytest.raises(sgraph.SgError) as e:
        srv.wait(srv.submit(sgraph.REQ_BFS, 0), sgraph.REQ_BFS, 0)
    assert e.value.status == sgraph.SG_EBUDGET
    out, info = srv.wait(srv.submit(sgraph.REQ_CC, 3), sgraph.REQ_CC, 3)
    assert out[0] == O.cc(row, col)[3]
    st = srv.stats()
    srv.close()
    assert st["deferred_for_memory"] >= 1 and st["queries"] == 0


def test_server_enters_failed_state_after_an_engine_error():
    # the smallest device budget the server can start with, then an update
    # batch whose scratch does not fit: that update fails with the engine's
    # status, and every later request fails with the saved error
    nv, row, col = sgraph.gen_rmat(12, 8, seed=103)
    srv, b = None, 0.001
    while srv is None and b < 1.0:
        try:
            srv = sgraph.Server(row, col, query_slots=4, standing=sgraph.SERVE_CC, device_budget_gb=b, hbm_reserve_gb=0)
        except sgraph.SgError as e:
            assert e.status in (sgraph.SG_EBUDGET, sgraph.SG_ENOMEM)
            b *= 1.25
    assert srv is not None
    rng = np.random.default_rng(104)
    big = rng.integers(0, nv, (400000, 2)).astype(np.uint32)
    with pytest.raises(sgraph.SgError) as e1:
        srv.wait(srv.submit(sgraph.REQ_UPDATE, ins=big, dele=big[:1000]))
    assert e1.value.status in (sgraph.SG_EBUDGET, sgraph.SG_ENOMEM)
    with pytest.raises(sgraph.SgError) as e2:
        srv.wait(srv.submit(sgraph.REQ_CC, 1), sgraph.REQ_CC, 1)
    assert e2.value.status == e1.value.status
    srv.close()


def test_one_waiter_per_ticket():
    # a second waiter on a ticket (while the first waits, or after it took the
    # result) gets SG_EINVAL; the first gets the answer
    nv, row, col = sgraph.gen_rmat(12, 8, seed=105)
    srv = sgraph.Server(row, col, query_slots=8, standing=sgraph.SERVE_CC)
    rng = np.random.default_rng(106)
    big = rng.integers(0, nv, (200000, 2)).astype(np.uint32)
    t = srv.submit(sgraph.REQ_UPDATE, ins=big, dele=big[:10])
    first = {}

    def waiter():
        try:
            first["out"] = srv.wait(t)
        except sgraph.SgError as e:
            first["err"] = e

    th = threading.Thread(target=waiter)
    th.start()
    import time
    time.sleep(0.05)
    with pytest.raises(sgraph.SgError) as e:
        srv.wait(t)
    assert e.value.status == sgraph.SG_EINVAL
    th.join()
    assert "out" in first and first["out"][1]["snapshot"] == 1
    q = srv.submit(sgraph.REQ_BFS, 0, target=1)
    srv.wait(q, sgraph.REQ_BFS, 0, 1)
    with pytest.raises(sgraph.SgError) as e:
        srv.wait(q, sgraph.REQ_BFS, 0, 1)
    assert e.value.status == sgraph.SG_EINVAL
    srv.close()
