"""A caller's own operator compiled into an engine (include/sgraph_operator.h):
widest path from a source, built outside the library as
libsg_widest_path.so and driven through the ordinary C ABI.

After every snapshot of an update stream the widths equal the oracle's
widest-first search (a textbook algorithm, pinned by brute force over every
simple path and by a fixed-point iteration) and the parents pass the widest-
path certificate (no edge widens a vertex, every reached vertex has a tight
parent edge, parent chains end at the source): resident, with the in-edge
store (recompute seeded from the in-lists, pull rounds in every mode), with
edges in host memory, and across a checkpoint round trip.
"""
import numpy as np
import pytest

import oracle as O
import sgraph

pytestmark = pytest.mark.gpu

W = dict(weight_mode="seeded", weight_seed=7)


def say(*a):
    print(*a, flush=True)


def widest_engine(**cfg):
    return sgraph.Engine("widest_path", create=sgraph.load_operator("widest_path"), **W, **cfg)


def check(e, r, c, src, tag):
    wd = e.values()
    exp = O.widest_path(r, c, src, "seeded", 7)
    assert np.array_equal(wd, exp), f"{tag}: widths differ from the oracle at {np.flatnonzero(wd != exp)[:5]}"
    par = e.gather(sgraph.Q_PARENT)
    assert O.widest_certificate(r, c, src, wd, par, "seeded", 7) == 0, f"{tag}: certificate"
    return wd


def widest_stream(scale, ef, nsnap, nupd, seed, **cfg):
    nv, row, col = sgraph.gen_rmat(scale, ef, seed=seed)
    ins, dele = sgraph.gen_updates(row, col, scale, nsnap, nupd, nupd, seed=seed + 1, hub_frac=0.5, hub_top=0.01)
    src = int(np.argmax(np.diff(row.astype(np.int64))))  # a hub: most vertices reached
    say(f"[gpu] widest path nv={nv} ne={len(col)} src={src} {cfg}")
    e = widest_engine(**cfg)
    try:
        e.load_csr(row, col)
        e.init(src)
        wd = check(e, row, col, src, "G0")
        assert (wd > 0).sum() > nv // 4
        r, c = row, col
        inval = 0
        for t in range(nsnap):
            e.apply(ins[t], dele[t])
            r, c = O.apply_updates(r, c, ins[t], dele[t])
            check(e, r, c, src, f"G{t + 1}")
            inval += e.stats()["invalidated"]
        st = e.stats()
        say(f"[gpu] widest path: invalidated {inval}, pull rounds {st['pull_rounds']}")
        assert inval > 0  # deletions cut tree edges: the decremental path ran
        return st
    finally:
        e.close()


def test_widest_path_resident():
    widest_stream(13, 16, 3, 3000, seed=71)


@pytest.mark.parametrize("div", [0, sgraph.PULL_ALWAYS, sgraph.PULL_NEVER])
def test_widest_path_in_edges_and_pull_rounds(div):
    st = widest_stream(13, 16, 3, 3000, seed=73, flags=sgraph.FLAG_IN_EDGES, pull_divisor=div)
    assert st["in_capacity"] > 0
    if div == sgraph.PULL_NEVER:
        assert st["pull_rounds"] == 0


def test_widest_path_host_edges():
    st = widest_stream(12, 16, 3, 2000, seed=75, flags=sgraph.FLAG_HOST_EDGES | sgraph.FLAG_IN_EDGES, segments=7,
                       cache_level=1, pull_divisor=sgraph.PULL_ALWAYS)
    assert st["host_edges"] == 1


def test_widest_path_checkpoint(tmp_path):
    nv, row, col = sgraph.gen_rmat(12, 16, seed=77)
    ins, dele = sgraph.gen_updates(row, col, 12, 2, 2000, 2000, seed=78)
    src = int(np.argmax(np.diff(row.astype(np.int64))))
    e = widest_engine()
    e.load_csr(row, col)
    e.init(src)
    e.apply(ins[0], dele[0])
    path = str(tmp_path / "widest.ckpt")
    e.save_state(path)
    e.close()
    f = widest_engine()
    f.restore_state(path)
    f.apply(ins[1], dele[1])
    r, c = O.apply_updates(*O.apply_updates(row, col, ins[0], dele[0]), ins[1], dele[1])
    check(f, r, c, src, "restored G2")
    f.close()
    # another program's checkpoint is refused
    g = sgraph.Engine("sssp", **W)
    with pytest.raises(sgraph.SgError):
        g.restore_state(path)
    g.close()
