"""ctypes binding of the CPU oracle (oracle.c). TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the checker; never by the product."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
INF = 0xFFFFFFFF
_lib = None
_P = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "or_weight": (C.c_uint32, [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
            "or_bfs_levels": (C.c_int, [C.c_uint32, _P, _P, C.c_uint32, _P]),
            "or_sssp_dijkstra": (C.c_int, [C.c_uint32, _P, _P, C.c_uint32, C.c_uint32, C.c_uint64, _P]),
            "or_sssp_fifo": (C.c_int, [C.c_uint32, _P, _P, C.c_uint32, C.c_uint32, C.c_uint64, _P]),
            "or_widest_path": (C.c_int, [C.c_uint32, _P, _P, C.c_uint32, C.c_uint32, C.c_uint64, _P]),
            "or_cc_min_ancestor": (C.c_int, [C.c_uint32, _P, _P, _P]),
            "or_cc_label_prop": (C.c_int, [C.c_uint32, _P, _P, _P]),
            "or_pr_fp64": (C.c_int, [C.c_uint32, _P, _P, C.c_double, C.c_double, C.c_int, _P]),
            "or_pr_host_f32": (C.c_int, [C.c_uint32, _P, _P, C.c_double, C.c_double, _P]),
            "or_apply_updates": (C.c_int, [C.c_uint32, _P, _P, _P, C.c_uint64, _P, C.c_uint64, C.POINTER(_P), C.POINTER(_P)]),
            "or_free": (None, [_P]),
            "or_pma_geometry": (None, [C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
            "or_segment_partition": (C.c_uint32, [C.c_uint32, _P, C.c_uint32, _P]),
            "or_sp_certificate": (C.c_uint64, [C.c_uint32, _P, _P, C.c_uint32, C.c_uint32, C.c_uint64, _P, _P]),
            "or_cc_certificate": (C.c_uint64, [C.c_uint32, _P, _P, _P, _P]),
            "or_pr_residual": (None, [C.c_uint32, _P, _P, C.c_double, _P, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)


def _csr(row, col):
    return np.ascontiguousarray(row, dtype=np.uint64), np.ascontiguousarray(col, dtype=np.uint32)


WMODE = {"hash128": 0, "seeded": 1, "one": 2, "hash128_noplus1": 3}


def weight(mode, seed, u, v):
    return lib().or_weight(WMODE.get(mode, mode), seed, u, v)


def bfs(row, col, src):
    row, col = _csr(row, col)
    out = np.empty(len(row) - 1, dtype=np.uint32)
    assert lib().or_bfs_levels(len(row) - 1, _p(row), _p(col), src, _p(out)) == 0
    return out


def sssp(row, col, src, wmode="hash128", seed=0):
    row, col = _csr(row, col)
    out = np.empty(len(row) - 1, dtype=np.uint32)
    assert lib().or_sssp_dijkstra(len(row) - 1, _p(row), _p(col), src, WMODE.get(wmode, wmode), seed, _p(out)) == 0
    return out


def sssp_fifo(row, col, src, wmode="hash128", seed=0):
    row, col = _csr(row, col)
    out = np.empty(len(row) - 1, dtype=np.uint32)
    assert lib().or_sssp_fifo(len(row) - 1, _p(row), _p(col), src, WMODE.get(wmode, wmode), seed, _p(out)) == 0
    return out


def widest_path(row, col, src, wmode="hash128", seed=0):
    """Widest (maximum-bottleneck) path widths from src: INF at src, 0 unreached."""
    row, col = _csr(row, col)
    out = np.empty(len(row) - 1, dtype=np.uint32)
    assert lib().or_widest_path(len(row) - 1, _p(row), _p(col), src, WMODE.get(wmode, wmode), seed, _p(out)) == 0
    return out


def edge_weights(row, col, wmode="hash128", seed=0):
    """Weight of every CSR edge (the engine's weight function)."""
    row, col = _csr(row, col)
    src = np.repeat(np.arange(len(row) - 1, dtype=np.uint32), np.diff(row.astype(np.int64)))
    w = np.fromiter((weight(wmode, seed, int(u), int(v)) for u, v in zip(src, col)), dtype=np.uint32, count=len(col))
    return src, w


def widest_certificate(row, col, src, width, parent, wmode="hash128", seed=0):
    """Violations of: width(src) = INF; no edge u -> v offers more than width(v)
    (min(width(u), w) <= width(v)); every reached v != src has a parent edge
    (p, v) with min(width(p), w(p, v)) = width(v), unreached vertices none; the
    parent chains are acyclic and end at src. With these, width(v) is the
    widest-path width (a path of that width exists, and no edge can widen it)."""
    row, col = _csr(row, col)
    nv = len(row) - 1
    width = np.asarray(width, dtype=np.uint64)
    parent = np.asarray(parent, dtype=np.uint64)
    s, w = edge_weights(row, col, wmode, seed)
    bad = int(width[src] != INF)
    offer = np.minimum(width[s], w.astype(np.uint64))
    bad += int(np.count_nonzero(offer > width[col]))
    reached = (width > 0) & (np.arange(nv) != src)
    bad += int(np.count_nonzero((width == 0) & (parent != INF)))
    rv = np.nonzero(reached)[0]
    p = parent[rv]
    okp = p < nv
    bad += int(np.count_nonzero(~okp))
    rv, p = rv[okp], p[okp]
    # the parent edge must exist with the tight weight (the widest of parallel edges)
    key_e = s.astype(np.uint64) << np.uint64(32) | col.astype(np.uint64)
    ukeys = np.unique(key_e)
    order = np.argsort(key_e, kind="stable")
    ks, ws = key_e[order], w[order]
    starts = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]]) if len(ks) else np.zeros(0, dtype=np.int64)
    wmax = np.maximum.reduceat(ws, starts) if len(ks) else np.zeros(0, dtype=np.uint32)
    want = p << np.uint64(32) | rv.astype(np.uint64)
    pos = np.searchsorted(ukeys, want)
    found = (pos < len(ukeys)) & (ukeys[np.minimum(pos, max(len(ukeys) - 1, 0))] == want) if len(ukeys) else \
        np.zeros(len(want), dtype=bool)
    bad += int(np.count_nonzero(~found))
    fi = np.nonzero(found)[0]
    tight = np.minimum(width[p[fi]], wmax[pos[fi]].astype(np.uint64)) == width[rv[fi]]
    bad += int(np.count_nonzero(~tight))
    # chains end at src: pointer jumping over the parent array
    nxt = np.where(reached & (parent < nv), parent, np.uint64(nv)).astype(np.int64)
    nxt[src] = src
    nxt = np.append(nxt, nv)  # sink for broken chains
    for _ in range(64):
        nxt = nxt[nxt]
    bad += int(np.count_nonzero(nxt[:nv][reached] != src))
    return bad


def cc(row, col):
    row, col = _csr(row, col)
    out = np.empty(len(row) - 1, dtype=np.uint32)
    assert lib().or_cc_min_ancestor(len(row) - 1, _p(row), _p(col), _p(out)) == 0
    return out


def cc_label_prop(row, col):
    row, col = _csr(row, col)
    out = np.empty(len(row) - 1, dtype=np.uint32)
    assert lib().or_cc_label_prop(len(row) - 1, _p(row), _p(col), _p(out)) == 0
    return out


def pagerank(row, col, alpha=0.85, tol=1e-13, max_iter=100000):
    row, col = _csr(row, col)
    out = np.empty(len(row) - 1, dtype=np.float64)
    it = lib().or_pr_fp64(len(row) - 1, _p(row), _p(col), alpha, tol, max_iter, _p(out))
    assert it > 0, "PageRank oracle did not converge"
    return out


def pagerank_host_f32(row, col, alpha=0.85, eps=1e-6):
    row, col = _csr(row, col)
    out = np.empty(len(row) - 1, dtype=np.float32)
    assert lib().or_pr_host_f32(len(row) - 1, _p(row), _p(col), alpha, eps, _p(out)) >= 0
    return out


def apply_updates(row, col, ins, dele):
    """Adjacency model of one snapshot: returns the new (row, col)."""
    row, col = _csr(row, col)
    ins = np.ascontiguousarray(ins, dtype=np.uint32).reshape(-1)
    dele = np.ascontiguousarray(dele, dtype=np.uint32).reshape(-1)
    nv = len(row) - 1
    pr, pc = C.c_void_p(), C.c_void_p()
    rc = lib().or_apply_updates(nv, _p(row), _p(col), _p(ins), len(ins) // 2, _p(dele), len(dele) // 2, C.byref(pr), C.byref(pc))
    assert rc == 0, f"or_apply_updates failed ({rc})"
    nrow = np.frombuffer((C.c_uint64 * (nv + 1)).from_address(pr.value), dtype=np.uint64).copy()
    ne = int(nrow[-1])
    ncol = np.frombuffer((C.c_uint32 * max(ne, 1)).from_address(pc.value), dtype=np.uint32)[:ne].copy()
    lib().or_free(pr)
    lib().or_free(pc)
    return nrow, ncol


def pma_geometry(nv):
    s, c = C.c_uint32(), C.c_uint64()
    lib().or_pma_geometry(nv, C.byref(s), C.byref(c))
    return s.value, c.value


def segment_partition(off, nseg):
    off = np.ascontiguousarray(off, dtype=np.uint64)
    out = np.zeros(nseg + 1, dtype=np.uint32)
    n = lib().or_segment_partition(len(off) - 1, _p(off), nseg, _p(out))
    return out[: n + 1]


def sp_certificate(row, col, src, dist, parent, wmode="hash128", seed=0):
    """Violations of the shortest-path certificate (0 = dist are the distances from src)."""
    row, col = _csr(row, col)
    dist = np.ascontiguousarray(dist, dtype=np.uint32)
    parent = np.ascontiguousarray(parent, dtype=np.uint32)
    return int(lib().or_sp_certificate(len(row) - 1, _p(row), _p(col), src, WMODE.get(wmode, wmode), seed, _p(dist),
                                       _p(parent)))


def cc_certificate(row, col, label, parent):
    """Violations of the min-ancestor label certificate (0 = label is the min-ancestor labelling)."""
    row, col = _csr(row, col)
    label = np.ascontiguousarray(label, dtype=np.uint32)
    parent = np.ascontiguousarray(parent, dtype=np.uint32)
    return int(lib().or_cc_certificate(len(row) - 1, _p(row), _p(col), _p(label), _p(parent)))


def pr_residual(row, col, rank, alpha=0.85):
    """(sum |res|, max |res|, sum rank) of the FP64 fixed-point residual of a rank vector.
    ||rank - r*||_1 <= sum|res| / (1 - alpha)."""
    row, col = _csr(row, col)
    rank = np.ascontiguousarray(rank, dtype=np.float32)
    l1, mx, sr = C.c_double(), C.c_double(), C.c_double()
    lib().or_pr_residual(len(row) - 1, _p(row), _p(col), alpha, _p(rank), C.byref(l1), C.byref(mx), C.byref(sr))
    return l1.value, mx.value, sr.value
