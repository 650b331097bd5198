"""sg_gather_at_noalloc (the lookup staged through the library's fixed device
scratch, no device allocation) returns what sg_gather_at returns, bit for bit,
on two calls in a row (the second with a smaller input, so stale scratch
contents from the first would show), for every kind of engine; sizes past its
maximum are refused with SG_EINVAL."""
import numpy as np
import pytest

import sgraph

pytestmark = pytest.mark.gpu

MAX = 65536


def say(*a):
    print(*a, flush=True)


def lookups(eng, nv, rng, q, qi=0):
    for n in (MAX, 1000):  # the full scratch, then a smaller call over the same scratch
        v = rng.integers(0, nv, n).astype(np.uint32)
        v[::97] = 0xFFFFFFFF  # ids >= nv read as UINT32_MAX by both entry points
        a = eng.gather_at_noalloc(v, q=q, qi=qi)
        b = eng.gather_at(v, q=q, qi=qi)
        assert np.array_equal(a, b), f"query {q}: {int((a != b).sum())} of {n} differ"


@pytest.mark.parametrize("algo", ["sssp", "cc", "pr"])
def test_noalloc_matches_gather_at(algo):
    say(f"[gpu] noalloc {algo}")
    nv, row, col = sgraph.gen_rmat(14, 8, seed=71)
    ins, dele = sgraph.gen_updates(row, col, 14, 1, 2000, 2000, seed=72)
    eng = sgraph.Engine(algo, weight_mode="seeded", weight_seed=5)
    eng.load_csr(row, col)
    eng.init(0 if algo == "sssp" else None)
    eng.apply(ins[0], dele[0])
    rng = np.random.default_rng(73)
    qs = [sgraph.Q_VALUE, sgraph.Q_DEGREE] + ([sgraph.Q_PARENT] if algo != "pr" else [sgraph.Q_BUFFER])
    for q in qs:
        lookups(eng, nv, rng, q)
    with pytest.raises(sgraph.SgError) as e:
        eng.gather_at_noalloc(np.zeros(MAX + 1, dtype=np.uint32))
    assert e.value.status == sgraph.SG_EINVAL
    eng.close()


def test_noalloc_batched_queries_and_deferred_pagerank():
    nv, row, col = sgraph.gen_rmat(13, 8, seed=74)
    eng = sgraph.Engine("multi", weight_mode="seeded", weight_seed=6)
    eng.load_csr(row, col)
    eng.init_queries([0, 5, 9], ["bfs", "sssp", "sssp"])
    rng = np.random.default_rng(75)
    for qi in range(3):
        lookups(eng, nv, rng, sgraph.Q_VALUE, qi=qi)
    eng.close()
    # PageRank with deferred snapshots: refused until they are converged
    ins, dele = sgraph.gen_updates(row, col, 13, 1, 500, 500, seed=76)
    pr = sgraph.Engine("pr", flags=sgraph.FLAG_DEFER_CONVERGE)
    pr.load_csr(row, col)
    pr.init()
    pr.apply(ins[0], dele[0])
    with pytest.raises(sgraph.SgError) as e:
        pr.gather_at_noalloc(np.arange(10, dtype=np.uint32))
    assert e.value.status == sgraph.SG_EINVAL
    pr.converge()
    lookups(pr, nv, rng, sgraph.Q_VALUE)
    pr.close()
