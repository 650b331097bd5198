"""The autotuner picks a point of its search space, and the engine run with
the picked configuration gives the oracle's results (candidates change only
scheduling)."""
import numpy as np
import pytest

import autotune
import oracle as O
import sgraph

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_candidate_with_correct_results():
    nv, row, col = sgraph.gen_rmat(11, 8, seed=21)
    ins, dele = sgraph.gen_updates(row, col, 11, 2, 300, 300, seed=22)
    space = {"wave_budget": [512, 2048], "rounds_per_sync": [2, 8], "graph_rounds": [False, True]}
    best, table = autotune.autotune("sssp", row, col, ins, dele, space=space, weight_mode="seeded", weight_seed=2)
    assert len(table) == 8
    assert best["wave_budget"] in (512, 2048) and best["rounds_per_sync"] in (2, 8)
    assert "graph_rounds" not in best and best.get("flags", 0) in (0, sgraph.FLAG_GRAPH_ROUNDS)
    e = sgraph.Engine("sssp", **best)
    e.load_csr(row, col)
    e.init(0)
    r, c = row, col
    for t in range(2):
        e.apply(ins[t], dele[t])
        r, c = O.apply_updates(r, c, ins[t], dele[t])
    assert np.array_equal(e.values(), O.sssp(r, c, 0, "seeded", 2))
    e.close()


def test_autotune_host_edges_space():
    nv, row, col = sgraph.gen_rmat(10, 8, seed=23)
    ins, dele = sgraph.gen_updates(row, col, 10, 2, 200, 200, seed=24)
    space = {"segments": [4, 32], "cache_level": [0, 1]}
    best, table = autotune.autotune("cc", row, col, ins, dele, space=space, flags=sgraph.FLAG_HOST_EDGES)
    assert len(table) == 4 and best["segments"] in (4, 32)
