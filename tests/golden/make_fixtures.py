"""Writes the golden fixtures of this directory. Expected outputs are computed
by brute force in pure Python (no oracle, no engine):

  readme_*.txt          the toy input of the reference's README.md:66-95
                        (graph, update file, stream-size file)
  readme_expected.json  per snapshot G_0..G_2: edge list in adjacency order,
                        BFS levels and SSSP distances from 0 (weight
                        (u+v)%128+1), min-ancestor labels, PageRank (exact
                        dense solve of r = (1-a) + a A^T D^-1 r)
  tiny_expected.json    seeded random multi-snapshot cases, same outputs plus
                        SSSP with the seeded [1,63] weight

BFS/SSSP use Bellman-Ford to a fixed point, labels use the transitive closure
(Floyd-Warshall reachability), PageRank uses numpy.linalg.solve. Run:
    python tests/golden/make_fixtures.py
"""
import json
import os
import random

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INF = 0xFFFFFFFF
MASK = (1 << 64) - 1

README_GRAPH = [(0, 1), (0, 2), (1, 2), (1, 3), (2, 4), (3, 4)]
README_UPDATES = [("a", 0, 3), ("a", 0, 4), ("a", 1, 4), ("a", 2, 3),
                  ("d", 0, 1), ("d", 0, 2), ("d", 1, 2), ("d", 1, 3), ("d", 2, 4), ("d", 3, 4)]
README_SIZES = [(2, 3), (2, 3)]


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK
    return x ^ (x >> 31)


def weight(mode, seed, u, v):
    if mode == "hash128":
        return (u + v) % 128 + 1
    if mode == "seeded":
        return 1 + splitmix64(seed ^ ((u << 32) | v)) % 63
    return 1


def apply(adj, ins, dele):
    """Inserts append to the source's list in order; then every entry equal to
    a deleted target is removed (order of the rest kept)."""
    adj = [list(a) for a in adj]
    for u, v in ins:
        adj[u].append(v)
    dset = set(dele)
    return [[v for v in adj[u] if (u, v) not in dset] for u in range(len(adj))]


def bellman_ford(adj, src, w):
    n = len(adj)
    d = [INF] * n
    d[src] = 0
    changed = True
    while changed:
        changed = False
        for u in range(n):
            if d[u] == INF:
                continue
            for v in adj[u]:
                nd = d[u] + w(u, v)
                if nd < d[v]:
                    d[v] = nd
                    changed = True
    return d


def min_ancestor(adj):
    n = len(adj)
    reach = [[False] * n for _ in range(n)]
    for u in range(n):
        reach[u][u] = True
        for v in adj[u]:
            reach[u][v] = True
    for k in range(n):
        for i in range(n):
            if reach[i][k]:
                for j in range(n):
                    if reach[k][j]:
                        reach[i][j] = True
    return [min(u for u in range(n) if reach[u][v]) for v in range(n)]


def pagerank(adj, alpha=0.85):
    n = len(adj)
    M = np.zeros((n, n))
    for u in range(n):
        if adj[u]:
            for v in adj[u]:
                M[v, u] += 1.0 / len(adj[u])
    return np.linalg.solve(np.eye(n) - alpha * M, np.full(n, 1.0 - alpha)).tolist()


def outputs(adj, src, seed):
    return {
        "adj": adj,
        "bfs": bellman_ford(adj, src, lambda u, v: 1),
        "sssp_hash128": bellman_ford(adj, src, lambda u, v: weight("hash128", 0, u, v)),
        "sssp_seeded": bellman_ford(adj, src, lambda u, v: weight("seeded", seed, u, v)),
        "cc": min_ancestor(adj),
        "pr": pagerank(adj),
    }


def readme():
    with open(os.path.join(HERE, "readme_graph.txt"), "w") as f:
        f.writelines(f"{u} {v}\n" for u, v in README_GRAPH)
    with open(os.path.join(HERE, "readme_update.txt"), "w") as f:
        f.writelines(f"{t} {u} {v}\n" for t, u, v in README_UPDATES)
    with open(os.path.join(HERE, "readme_stream_size.txt"), "w") as f:
        f.write("2 3 # \"2\" is the number of added edges, \"3\" is the number of deleted edges.\n2 3\n")
    n = 5
    adj = [[] for _ in range(n)]
    for u, v in README_GRAPH:
        adj[u].append(v)
    ins = [(u, v) for t, u, v in README_UPDATES if t == "a"]
    dele = [(u, v) for t, u, v in README_UPDATES if t == "d"]
    snaps = [outputs(adj, 0, 27491095)]
    io = do = 0
    for na, nd in README_SIZES:
        adj = apply(adj, ins[io:io + na], dele[do:do + nd])
        io += na
        do += nd
        snaps.append(outputs(adj, 0, 27491095))
    with open(os.path.join(HERE, "readme_expected.json"), "w") as f:
        json.dump({"nv": n, "source": 0, "seed": 27491095, "snapshots": snaps}, f, indent=1)


def tiny():
    rng = random.Random(27491095)
    cases = []
    for case in range(6):
        n = rng.randint(6, 24)
        m = rng.randint(n, 4 * n)
        edges = set()
        while len(edges) < m:
            u, v = rng.randrange(n), rng.randrange(n)
            if u != v:
                edges.add((u, v))
        g = sorted(edges)
        adj = [[] for _ in range(n)]
        for u, v in g:
            adj[u].append(v)
        snaps_in = []
        cur = set(g)
        for _ in range(3):
            cand = [(u, v) for u in range(n) for v in range(n) if u != v and (u, v) not in cur]
            ins = rng.sample(cand, min(len(cand), rng.randint(0, n)))
            dele = rng.sample(sorted(cur), min(len(cur), rng.randint(0, n)))
            if rng.random() < 0.5 and ins:   # an edge inserted and deleted in one snapshot
                dele.append(ins[0])
            snaps_in.append({"ins": ins, "del": dele})
            cur = (cur | set(ins)) - set(dele)
        src = rng.randrange(n)
        out = [outputs(adj, src, 12345)]
        for s in snaps_in:
            adj = apply(adj, s["ins"], s["del"])
            out.append(outputs(adj, src, 12345))
        cases.append({"nv": n, "source": src, "seed": 12345, "edges": g, "updates": snaps_in, "snapshots": out})
    with open(os.path.join(HERE, "tiny_expected.json"), "w") as f:
        json.dump(cases, f)


if __name__ == "__main__":
    readme()
    tiny()
