"""Input formats and synthetic inputs of libsgraph.so (host code, no GPU)."""
import os

import numpy as np
import pytest

import oracle as O
import sgraph
from conftest import GOLDEN


def test_readme_files_parse():
    nv, row, col = sgraph.read_edge_list(os.path.join(GOLDEN, "readme_graph.txt"))
    assert nv == 5
    assert row.tolist() == [0, 2, 4, 5, 6, 6]
    assert col.tolist() == [1, 2, 2, 3, 4, 4]
    sizes = sgraph.read_stream_sizes(os.path.join(GOLDEN, "readme_stream_size.txt"))
    assert sizes.tolist() == [[2, 3], [2, 3]]
    ins, dele = sgraph.read_updates(os.path.join(GOLDEN, "readme_update.txt"))
    assert ins.tolist() == [[0, 3], [0, 4], [1, 4], [2, 3]]
    assert dele.tolist() == [[0, 1], [0, 2], [1, 2], [1, 3], [2, 4], [3, 4]]


def test_edge_list_semantics(tmp_path):
    p = tmp_path / "g.txt"
    # file order kept per source, duplicates and self-loops kept, weights ignored,
    # junk lines skipped, nv from max(src, dst) + 1
    p.write_text("2 1\n0 5 17\n# comment\n2 0\n0 5\n\n3 3\n")
    nv, row, col = sgraph.read_edge_list(str(p))
    assert nv == 6
    assert row.tolist() == [0, 2, 2, 4, 5, 5, 5]
    assert col.tolist() == [5, 5, 1, 0, 3]


def test_update_file_semantics(tmp_path):
    p = tmp_path / "u.txt"
    p.write_text("d 1 2 9\na 0 1\nx 5 5\nd 3 4\na 2 2 7\nab 1 1\n")
    ins, dele = sgraph.read_updates(str(p))
    assert ins.tolist() == [[0, 1], [2, 2]]
    assert dele.tolist() == [[1, 2], [3, 4]]


def test_round_trip(tmp_path):
    nv, row, col = sgraph.gen_rmat(8, 4)
    g = str(tmp_path / "g.txt")
    sgraph.write_edge_list(g, row, col)
    nv2, row2, col2 = sgraph.read_edge_list(g)
    # nv of a file is max id + 1
    assert np.array_equal(row2, row[: nv2 + 1]) and np.array_equal(col2, col)
    ins, dele = sgraph.gen_updates(row, col, 8, 3, 5, 4)
    sizes = np.array([[5, 4]] * 3, dtype=np.uint64)
    u, s = str(tmp_path / "u.txt"), str(tmp_path / "s.txt")
    sgraph.write_updates(u, s, sizes, ins, dele)
    i2, d2 = sgraph.read_updates(u)
    assert np.array_equal(i2, ins.reshape(-1, 2)) and np.array_equal(d2, dele.reshape(-1, 2))
    assert np.array_equal(sgraph.read_stream_sizes(s), sizes)


def test_missing_file_is_eio():
    with pytest.raises(sgraph.SgError) as e:
        sgraph.read_edge_list("/nonexistent/graph.txt")
    assert e.value.status == sgraph.SG_EIO


def test_rmat_properties():
    nv, row, col = sgraph.gen_rmat(12, 16)
    assert nv == 4096
    nv2, row2, col2 = sgraph.gen_rmat(12, 16)
    assert np.array_equal(row, row2) and np.array_equal(col, col2)
    ne = len(col)
    assert 0.5 * 16 * nv < ne <= 16 * nv
    src = np.repeat(np.arange(nv, dtype=np.uint64), np.diff(row).astype(np.int64))
    keys = (src << np.uint64(32)) | col.astype(np.uint64)
    assert np.all(np.diff(keys.astype(np.int64)) > 0)  # sorted, no duplicates
    assert not np.any(src == col)  # no self-loops
    deg = np.diff(row)
    assert deg[0] == deg.max()  # unpermuted R-MAT: vertex 0 is the hub
    _, rowp, colp = sgraph.gen_rmat(12, 16, permute=True)
    assert len(colp) == ne  # permutation preserves the edge count


def test_update_stream_properties():
    nv, row, col = sgraph.gen_rmat(10, 8)
    ins, dele = sgraph.gen_updates(row, col, 10, 4, 300, 200, seed=5)
    cur = set(zip(np.repeat(np.arange(nv), np.diff(row).astype(np.int64)).tolist(), col.tolist()))
    for t in range(4):
        batch_i = [tuple(x) for x in ins[t].tolist()]
        batch_d = [tuple(x) for x in dele[t].tolist()]
        assert len(set(batch_i)) == len(batch_i) and len(set(batch_d)) == len(batch_d)
        for e in batch_i:
            assert e not in cur and e[0] != e[1]
        for e in batch_d:
            assert e in cur  # deletes come from G_{t-1}
        cur = (cur | set(batch_i)) - set(batch_d)
    # hub bias: most inserted sources among the top 1% by out-degree
    ins_h, _ = sgraph.gen_updates(row, col, 10, 1, 500, 0, seed=6, hub_frac=0.8, hub_top=0.01)
    deg = np.diff(row)
    k = max(1, int(0.01 * nv))
    hubs = set(np.argsort(-deg.astype(np.int64), kind="stable")[:k].tolist())
    frac = np.mean([u in hubs for u in ins_h[0, :, 0].tolist()])
    assert frac > 0.7


def test_weights_match_oracle():
    rng = np.random.default_rng(0)
    for _ in range(2000):
        u, v = int(rng.integers(0, 2**32 - 1)), int(rng.integers(0, 2**32 - 1))
        for m in ("hash128", "seeded", "one"):
            w = sgraph.edge_weight(m, 1234567, u, v)
            assert w == O.weight(m, 1234567, u, v)
            assert 1 <= w <= 128
        assert sgraph.edge_weight("seeded", 99, u, v) <= 63


def test_snapshot_file_round_trip_and_checks(tmp_path):
    # the binary snapshot file holds exactly what the text files hold: the
    # README fixture's graph and stream read back equal, and a truncated or
    # altered file is refused (SG_EIO), never half-read
    nv, row, col = sgraph.gen_rmat(9, 8, seed=5)
    ins, dele = sgraph.gen_updates(row, col, 9, 3, 70, 40, seed=6)
    sizes = np.array([[70, 40]] * 3, dtype=np.uint64)
    p = str(tmp_path / "g.sgsnap")
    sgraph.write_snapshot_file(p, row, col, sizes, np.concatenate(ins), np.concatenate(dele))
    r2, c2, s2, i2, d2 = sgraph.read_snapshot_file(p)
    assert np.array_equal(r2, row) and np.array_equal(c2, col) and np.array_equal(s2, sizes)
    assert np.array_equal(i2, np.concatenate(ins).reshape(-1, 2)) and np.array_equal(d2, np.concatenate(dele).reshape(-1, 2))
    # snapshot t of the file applies like the generator's snapshot t
    r, c = O.apply_updates(r2, c2, i2[70:140], d2[40:80])
    r0, c0 = O.apply_updates(row, col, ins[1], dele[1])
    assert np.array_equal(r, r0) and np.array_equal(c, c0)
    raw = open(p, "rb").read()
    for bad in (raw[:-4], raw + b"\0", b"SGSNAP02" + raw[8:]):
        q = str(tmp_path / "bad.sgsnap")
        open(q, "wb").write(bad)
        with pytest.raises(sgraph.SgError) as e:
            sgraph.read_snapshot_file(q)
        assert e.value.status == sgraph.SG_EIO
    # offsets that do not end at ne
    hdr = 8 + 8 + 24
    b = bytearray(raw)
    b[hdr + 8 * nv: hdr + 8 * nv + 8] = (int(row[-1]) - 1).to_bytes(8, "little")
    q = str(tmp_path / "bad2.sgsnap")
    open(q, "wb").write(bytes(b))
    with pytest.raises(sgraph.SgError):
        sgraph.read_snapshot_file(q)
    # the README fixture's three text files as one snapshot file
    _, gr, gc = sgraph.read_edge_list(os.path.join(GOLDEN, "readme_graph.txt"))
    gi, gd = sgraph.read_updates(os.path.join(GOLDEN, "readme_update.txt"))
    gs = sgraph.read_stream_sizes(os.path.join(GOLDEN, "readme_stream_size.txt"))
    q = str(tmp_path / "readme.sgsnap")
    sgraph.write_snapshot_file(q, gr, gc, gs, gi, gd)
    r3, c3, s3, i3, d3 = sgraph.read_snapshot_file(q)
    assert np.array_equal(r3, gr) and np.array_equal(c3, gc) and np.array_equal(s3.reshape(-1), np.asarray(gs).reshape(-1))
    assert np.array_equal(i3, gi) and np.array_equal(d3, gd)


def test_snapshot_file_crafted_counts_are_refused(tmp_path):
    # header counts whose byte sizes wrap 64 bits (so the wrapped total equals
    # the file's real length) and stream sizes whose sum wraps are refused
    # before anything is allocated or read: SG_EIO, never a short buffer
    nv, row, col = sgraph.gen_rmat(8, 8, seed=11)
    ins, dele = sgraph.gen_updates(row, col, 8, 3, 70, 40, seed=12)
    sizes = np.array([[70, 40]] * 3, dtype=np.uint64)
    p = str(tmp_path / "g.sgsnap")
    sgraph.write_snapshot_file(p, row, col, sizes, np.concatenate(ins), np.concatenate(dele))
    raw = bytearray(open(p, "rb").read())
    ne, ti, td = len(col), 210, 120
    h64 = 8 + 8  # magic, nv, nsnap; then ne, ti, td (u64 each)

    def refused(b, name):
        q = str(tmp_path / f"{name}.sgsnap")
        open(q, "wb").write(bytes(b))
        with pytest.raises(sgraph.SgError) as e:
            sgraph.read_snapshot_file(q)
        assert e.value.status == sgraph.SG_EIO, name

    sz = h64 + 24 + 8 * (nv + 1) + 4 * ne  # stream sizes [n_add, n_del] per snapshot
    b = bytearray(raw)  # ti * 8 wraps to the real ti * 8, and the sizes still sum to ti
    b[h64 + 8:h64 + 16] = ((1 << 61) + ti).to_bytes(8, "little")
    b[sz:sz + 8] = ((1 << 61) + 70).to_bytes(8, "little")
    refused(b, "wrap_ti")
    b = bytearray(raw)  # td * 8 wraps
    b[h64 + 16:h64 + 24] = ((1 << 61) + td).to_bytes(8, "little")
    b[sz + 8:sz + 16] = ((1 << 61) + 40).to_bytes(8, "little")
    refused(b, "wrap_td")
    b = bytearray(raw)  # ne * 4 wraps (and row[nv] claims it)
    b[h64:h64 + 8] = ((1 << 62) + ne).to_bytes(8, "little")
    b[h64 + 24 + 8 * nv:h64 + 24 + 8 * nv + 8] = ((1 << 62) + ne).to_bytes(8, "little")
    refused(b, "wrap_ne")
    # stream sizes whose sum wraps to the insert total
    b = bytearray(raw)
    b[sz:sz + 8] = ((1 << 64) - 100).to_bytes(8, "little")
    b[sz + 16:sz + 24] = (70 + 100 + 70).to_bytes(8, "little")
    refused(b, "wrap_sizes")


def test_splitmix64_known_answers():
    # The seeded weight is 1 + splitmix64(seed ^ (u << 32 | v)) % 63, where
    # splitmix64(x) is the first output of Vigna's splitmix64 generator seeded
    # with x. Published outputs of that generator seeded with 1234567 (Rosetta
    # Code, "Pseudo-random numbers/Splitmix64"): the k-th output is
    # splitmix64(1234567 + (k - 1) * 0x9E3779B97F4A7C15 mod 2^64).
    published = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                 4593380528125082431, 16408922859458223821]
    golden = 0x9E3779B97F4A7C15
    for k, z in enumerate(published):
        x = (1234567 + k * golden) % (1 << 64)
        want = 1 + z % 63
        # seed = x, (u, v) = (0, 0); and seed = 0 with x spread over (u, v)
        assert sgraph.edge_weight("seeded", x, 0, 0) == want
        assert sgraph.edge_weight("seeded", 0, x >> 32, x & 0xFFFFFFFF) == want
        assert O.weight("seeded", x, 0, 0) == want
