"""hybrid_pr / hybrid_sssp / hybrid_bfs / hybrid_cc command lines: the
reference's flags and input files (README.md:46-53 command), result files in
the reference's output formats, per-snapshot dumps checked against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, PKG

BIN = os.path.join(PKG, "bin")
G = os.path.join(GOLDEN, "readme_graph.txt")
U = os.path.join(GOLDEN, "readme_update.txt")
S = os.path.join(GOLDEN, "readme_stream_size.txt")
README_ARGS = ["-graphfile", G, "-format", "market_big", "-hybrid", "0", "-SEGMENT", "512", "-weight_num", "1",
               "-weight", "1", "-update_size", S, "-updatefile", U, "-cache", "0"]


def run(name, *args):
    return subprocess.run([os.path.join(BIN, name), *args], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("name", ["hybrid_pr", "hybrid_sssp", "hybrid_bfs", "hybrid_cc"])
def test_binaries_exist(name):
    assert os.access(os.path.join(BIN, name), os.X_OK)


def test_unknown_flag_and_missing_input_exit_1():
    r = run("hybrid_sssp", "-graphfile", G, "-no_such_flag", "3")
    assert r.returncode == 1 and "unknown command line flag" in r.stderr
    assert run("hybrid_sssp").returncode == 1
    assert run("hybrid_sssp", "-graphfile", "/nonexistent").returncode == 1
    assert run("hybrid_pr", "-graphfile", G, "-error", "abc").returncode == 1
    assert run("hybrid_pr", "-graphfile", G, "-format", "metis").returncode == 1


def test_help_lists_reference_flags():
    r = run("hybrid_sssp", "-help")
    assert r.returncode == 0
    for f in ("graphfile", "updatefile", "update_size", "SEGMENT", "n_stream", "cache", "source_node", "error"):
        assert f"-{f}:" in r.stdout


def read_result(path):
    out = []
    for line in open(path):
        i, v = line.split()
        out.append(0xFFFFFFFF if v == "INF" else int(v))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name,key", [("hybrid_sssp", "sssp_hash128"), ("hybrid_bfs", "bfs"), ("hybrid_cc", "cc")])
def test_readme_command_outputs(tmp_path, name, key):
    exp = json.load(open(os.path.join(GOLDEN, "readme_expected.json")))["snapshots"]
    out = str(tmp_path / "out.txt")
    r = run(name, *README_ARGS, "-source_node", "0", "-output", out, "-dump_dir", str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "Overall: Test passed" in r.stdout
    assert read_result(out) == exp[2][key]
    assert read_result(str(tmp_path / "snapshot_0.txt")) == exp[1][key]


@pytest.mark.gpu
def test_pr_command_output(tmp_path):
    exp = np.array(json.load(open(os.path.join(GOLDEN, "readme_expected.json")))["snapshots"][2]["pr"])
    out = str(tmp_path / "pr.txt")
    r = run("hybrid_pr", *README_ARGS, "-output", out, "-print_ranks")
    assert r.returncode == 0, r.stderr
    lines = open(out).read().splitlines()
    assert lines[0].startswith("ALPHA") and lines[1] == "RANKS 1--5 of 5"
    got = {int(a): float(b) for a, b in (l.split() for l in lines[2:])}
    np.testing.assert_allclose([got[i] for i in range(5)], exp, rtol=1e-4)
    ranks = [int(l.split()[0]) for l in lines[2:]]
    assert [got[v] for v in ranks] == sorted(got.values(), reverse=True)


@pytest.mark.gpu
def test_rmat_files_seeded_weights(tmp_path):
    import sgraph
    nv, row, col = sgraph.gen_rmat(10, 8, seed=81)
    g, u, s = str(tmp_path / "g.txt"), str(tmp_path / "u.txt"), str(tmp_path / "s.txt")
    sgraph.write_edge_list(g, row, col)
    nvf, rowf, colf = sgraph.read_edge_list(g)  # the file's vertex count is its max id + 1
    ins, dele = sgraph.gen_updates(rowf, colf, 9, 3, 100, 100, seed=82)  # ids < 512 <= nvf
    sgraph.write_updates(u, s, np.array([[100, 100]] * 3), ins, dele)
    r = run("hybrid_sssp", "-graphfile", g, "-updatefile", u, "-update_size", s, "-weight_mode", "seeded",
            "-weight_seed", "7", "-source_node", "3", "-dump_dir", str(tmp_path))
    assert r.returncode == 0, r.stderr
    rr, cc = rowf, colf
    for t in range(3):
        rr, cc = O.apply_updates(rr, cc, ins[t], dele[t])
        assert read_result(str(tmp_path / f"snapshot_{t}.txt")) == O.sssp(rr, cc, 3, "seeded", 7).tolist()


def test_write_snapshot_file_converts_text_inputs(tmp_path):
    # the conversion happens before any device call, so it runs without a GPU
    import sgraph
    snap = str(tmp_path / "readme.sgsnap")
    run("hybrid_bfs", *README_ARGS, "-write_snapshot_file", snap)
    row, col, sizes, ins, dele = sgraph.read_snapshot_file(snap)
    _, gr, gc = sgraph.read_edge_list(G)
    gi, gd = sgraph.read_updates(U)
    gs = sgraph.read_stream_sizes(S)
    assert np.array_equal(row, gr) and np.array_equal(col, gc) and np.array_equal(sizes, gs)
    assert np.array_equal(ins, gi[: int(gs[:, 0].sum())]) and np.array_equal(dele, gd[: int(gs[:, 1].sum())])
    r = run("hybrid_bfs", "-snapshot_file", str(tmp_path / "missing.sgsnap"))
    assert r.returncode == 1 and "snapshot file" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name,key", [("hybrid_sssp", "sssp_hash128"), ("hybrid_cc", "cc")])
def test_snapshot_file_run_equals_text_run(tmp_path, name, key):
    exp = json.load(open(os.path.join(GOLDEN, "readme_expected.json")))["snapshots"]
    snap = str(tmp_path / "readme.sgsnap")
    out = str(tmp_path / "text.txt")
    r = run(name, *README_ARGS, "-source_node", "0", "-output", out, "-write_snapshot_file", snap)
    assert r.returncode == 0, r.stderr
    out2 = str(tmp_path / "bin.txt")
    r = run(name, "-snapshot_file", snap, "-source_node", "0", "-output", out2, "-dump_dir", str(tmp_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert read_result(out2) == read_result(out) == exp[2][key]
    assert read_result(str(tmp_path / "snapshot_0.txt")) == exp[1][key]
