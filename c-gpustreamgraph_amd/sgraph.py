"""ctypes binding of libsgraph.so (include/sgraph.h, include/sgraph_io.h).

Host-side harness used by the tests and bench.py. Arrays cross the boundary
as numpy arrays; every call raises SgError on a non-OK status. The library is
loaded from this directory; a missing library raises instead of falling back
to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgraph.so")

SG_OK, SG_EINVAL, SG_ENOMEM, SG_EBUDGET, SG_EDEVICE, SG_EIO = range(6)
STATUS_NAMES = {0: "SG_OK", 1: "SG_EINVAL", 2: "SG_ENOMEM", 3: "SG_EBUDGET", 4: "SG_EDEVICE", 5: "SG_EIO"}
ALGO = {"pr": 0, "sssp": 1, "bfs": 2, "cc": 3, "multi": 4, "prcc": 5, "custom": 6}
WEIGHT = {"hash128": 0, "seeded": 1, "one": 2}
Q_VALUE, Q_BUFFER, Q_PARENT, Q_DEGREE, Q_OFFSETS, Q_SLOTS, Q_SEGMENTS, Q_IN_DEGREE, Q_IN_OFFSETS, Q_IN_SLOTS = range(10)
PULL_NEVER = 0xFFFFFFFF      # pull_divisor: push rounds only, no in-edge store
PULL_ALWAYS = 0x7FFFFFFF     # pull_divisor: every round with a non-empty frontier pulls
FLAG_KERNEL_TIMING = 0x1
FLAG_HOST_EDGES = 0x2
FLAG_DEFER_CONVERGE = 0x4
FLAG_GRAPH_ROUNDS = 0x8
FLAG_TIME_SWEEP = 0x10
FLAG_IN_EDGES = 0x20
FLAG_PULL_STREAM = 0x40
KCLASS = ("relax", "sweep", "merge", "pull")
INF = 0xFFFFFFFF
DEFAULT_SEED = 27491095  # kRandSeed of the reference (groute/graphs/util.h:68)


class SgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Config(C.Structure):
    _fields_ = [
        ("algo", C.c_uint32), ("weight_mode", C.c_uint32), ("weight_seed", C.c_uint64),
        ("pr_alpha", C.c_double), ("pr_epsilon", C.c_double), ("pma_density", C.c_double),
        ("segments", C.c_uint32), ("n_stream", C.c_uint32), ("pull_divisor", C.c_uint32),
        ("max_rounds", C.c_uint32), ("rounds_per_sync", C.c_uint32), ("flags", C.c_uint32),
        ("device_budget_gb", C.c_double), ("device", C.c_int32),
        ("wave_budget", C.c_uint32), ("heavy_degree", C.c_uint32), ("cache_level", C.c_uint32),
        ("edge_budget_gb", C.c_double), ("round_blocks", C.c_uint32), ("reserved", C.c_uint32 * 1),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("nv", C.c_uint32), ("rounds", C.c_uint32), ("ne", C.c_uint64), ("capacity", C.c_uint64),
        ("snapshots", C.c_uint64), ("edges_relaxed", C.c_uint64), ("vertices_processed", C.c_uint64),
        ("invalidated", C.c_uint64), ("rebalances", C.c_uint64), ("resizes", C.c_uint64),
        ("dropped_updates", C.c_uint64), ("relax_launches", C.c_uint64), ("relax_ms", C.c_double),
        ("relax_bytes", C.c_uint64), ("last_phase_ms", C.c_double), ("pr_residual_l1", C.c_double),
        ("pr_residual_max", C.c_double), ("device_bytes", C.c_uint64), ("host_edges", C.c_uint64),
        ("cache_slots", C.c_uint64), ("cached_vertices", C.c_uint64), ("edges_from_cache", C.c_uint64),
        ("streamed_bytes", C.c_uint64), ("stream_ms", C.c_double), ("host_bytes", C.c_uint64),
        ("cache_fill_bytes", C.c_uint64), ("cache_fill_ms", C.c_double), ("pull_rounds", C.c_uint64),
        ("in_capacity", C.c_uint64), ("in_mirror_slots", C.c_uint64), ("in_mirror_live", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ServerConfig(C.Structure):
    _fields_ = [("engine", Config), ("query_slots", C.c_uint32), ("standing", C.c_uint32),
                ("hbm_reserve_gb", C.c_double), ("reserved", C.c_uint32 * 4)]


class ResultInfo(C.Structure):
    _fields_ = [("snapshot", C.c_uint64), ("latency_ms", C.c_double), ("step", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64), ("snapshots", C.c_uint64), ("queries", C.c_uint64), ("lookups", C.c_uint64),
        ("edge_updates", C.c_uint64), ("edges_answered", C.c_uint64), ("max_batch", C.c_uint64),
        ("deferred_for_memory", C.c_uint64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
        ("query_p50_ms", C.c_double), ("query_p99_ms", C.c_double), ("update_p50_ms", C.c_double),
        ("update_p99_ms", C.c_double), ("busy_ms", C.c_double), ("wall_ms", C.c_double),
        ("reserved", C.c_uint64 * 4),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


REQ_UPDATE, REQ_BFS, REQ_SSSP, REQ_CC, REQ_PR_TOPK = range(5)
SERVE_CC, SERVE_PR = 1, 2

_P = C.c_void_p
_lib = None

# (name, restype, argtypes) of every exported function
SIGNATURES = [
    ("sg_config_default", None, [C.POINTER(Config)]),
    ("sg_create", C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    ("sg_destroy", None, [_P]),
    ("sg_last_error", C.c_char_p, [_P]),
    ("sg_load_csr", C.c_int, [_P, C.c_uint32, C.c_uint64, _P, _P]),
    ("sg_init", C.c_int, [_P, _P, C.c_uint32]),
    ("sg_apply_snapshot", C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint64]),
    ("sg_apply_snapshot_device", C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint64]),
    ("sg_gather", C.c_int, [_P, C.c_uint32, _P, C.c_size_t]),
    ("sg_init_queries", C.c_int, [_P, _P, _P, C.c_uint32]),
    ("sg_gather_at", C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint64, _P]),
    ("sg_gather_at_noalloc", C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint64, _P]),
    ("sg_pr_topk", C.c_int, [_P, C.c_uint32, _P, _P]),
    ("sg_converge", C.c_int, [_P]),
    ("sg_device_memory", C.c_int, [C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("sg_server_config_default", None, [C.POINTER(ServerConfig)]),
    ("sg_server_create", C.c_int, [C.POINTER(ServerConfig), C.c_uint32, C.c_uint64, _P, _P, C.POINTER(_P)]),
    ("sg_server_submit", C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_uint64, _P, C.c_uint64,
                                   C.POINTER(C.c_uint64)]),
    ("sg_server_wait", C.c_int, [_P, C.c_uint64, _P, C.c_size_t, C.POINTER(ResultInfo)]),
    ("sg_server_get_stats", C.c_int, [_P, C.POINTER(ServerStats)]),
    ("sg_server_last_error", C.c_char_p, [_P]),
    ("sg_server_destroy", None, [_P]),
    ("sg_gather_query", C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_size_t]),
    ("sg_query_open", C.c_int, [_P, C.c_uint32]),
    ("sg_query_add", C.c_int, [_P, _P, _P, C.c_uint32, _P]),
    ("sg_query_step", C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("sg_query_release", C.c_int, [_P, _P, C.c_uint32]),
    ("sg_get_stats", C.c_int, [_P, C.POINTER(Stats)]),
    ("sg_kernel_times", C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    ("sg_reset_stats", C.c_int, [_P]),
    ("sg_device_alloc", C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    ("sg_device_free", C.c_int, [_P, _P]),
    ("sg_copy_to_device", C.c_int, [_P, _P, _P, C.c_size_t]),
    ("sg_synchronize", C.c_int, [_P]),
    ("sg_free", None, [_P]),
    ("sg_read_edge_list", C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(_P), C.POINTER(_P)]),
    ("sg_write_edge_list", C.c_int, [C.c_char_p, C.c_uint32, _P, _P]),
    ("sg_read_stream_sizes", C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(_P)]),
    ("sg_read_updates", C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(_P), C.POINTER(C.c_uint64), C.POINTER(_P)]),
    ("sg_write_updates", C.c_int, [C.c_char_p, C.c_char_p, C.c_uint32, _P, _P, _P]),
    ("sg_gen_rmat_device", C.c_int, [C.c_int32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(_P), C.POINTER(_P)]),
    ("sg_save_state", C.c_int, [_P, C.c_char_p]),
    ("sg_restore_state", C.c_int, [_P, C.c_char_p]),
    ("sg_write_snapshot_file", C.c_int, [C.c_char_p, C.c_uint32, C.c_uint64, _P, _P, C.c_uint32, _P, _P, _P]),
    ("sg_read_snapshot_file", C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(_P),
                                        C.POINTER(_P), C.POINTER(C.c_uint32), C.POINTER(_P), C.POINTER(_P),
                                        C.POINTER(_P)]),
    ("sg_gen_rmat", C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_int,
                              C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(_P), C.POINTER(_P)]),
    ("sg_gen_updates", C.c_int, [C.c_uint32, _P, _P, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint32,
                                 C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_uint64, C.POINTER(_P), C.POINTER(_P)]),
    ("sg_edge_weight", C.c_uint32, [C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
]


def lib():
    """Load libsgraph.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_operators = {}


def load_operator(name, path=None):
    """Create function of an operator library built with include/sgraph_operator.h:
    `sg_create_<name>` from `path` (default: lib<...> beside libsgraph.so,
    libsg_<name>.so). Pass it to Engine(create=...)."""
    key = (name, path)
    if key not in _operators:
        lib()  # libsgraph.so first: the operator library links to it
        L = C.CDLL(path or os.path.join(HERE, f"libsg_{name}.so"))
        f = getattr(L, f"sg_create_{name}")
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_void_p]
        _operators[key] = (L, f)
    return _operators[key][1]


def _check(st, eng=None):
    if st != SG_OK:
        msg = lib().sg_last_error(eng)
        raise SgError(st, msg.decode() if msg else "")


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else C.c_void_p(0)


def _take(p, count, dtype):
    """Copy a malloc'd array returned by the library into numpy and free it."""
    if count == 0:
        out = np.zeros(0, dtype=dtype)
    else:  # one copy (arrays of billions of entries)
        buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(p.value)
        out = np.frombuffer(buf, dtype=dtype).copy()
    lib().sg_free(p)
    return out


# ---- formats and generators (no GPU) --------------------------------------
def read_edge_list(path):
    nv, ne, row, col = C.c_uint32(), C.c_uint64(), C.c_void_p(), C.c_void_p()
    _check(lib().sg_read_edge_list(path.encode(), C.byref(nv), C.byref(ne), C.byref(row), C.byref(col)))
    return nv.value, _take(row, nv.value + 1, np.uint64), _take(col, ne.value, np.uint32)


def write_edge_list(path, row, col):
    row = np.ascontiguousarray(row, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    _check(lib().sg_write_edge_list(path.encode(), len(row) - 1, _ptr(row), _ptr(col)))


def read_stream_sizes(path):
    n, p = C.c_uint32(), C.c_void_p()
    _check(lib().sg_read_stream_sizes(path.encode(), C.byref(n), C.byref(p)))
    return _take(p, 2 * n.value, np.uint64).reshape(-1, 2)


def read_updates(path):
    ni, pi, nd, pd = C.c_uint64(), C.c_void_p(), C.c_uint64(), C.c_void_p()
    _check(lib().sg_read_updates(path.encode(), C.byref(ni), C.byref(pi), C.byref(nd), C.byref(pd)))
    return _take(pi, 2 * ni.value, np.uint32).reshape(-1, 2), _take(pd, 2 * nd.value, np.uint32).reshape(-1, 2)


def write_updates(update_path, size_path, sizes, ins, dele):
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64).reshape(-1)
    ins = np.ascontiguousarray(ins, dtype=np.uint32).reshape(-1)
    dele = np.ascontiguousarray(dele, dtype=np.uint32).reshape(-1)
    _check(lib().sg_write_updates(update_path.encode(), size_path.encode(), len(sizes) // 2, _ptr(sizes), _ptr(ins), _ptr(dele)))


def write_snapshot_file(path, row, col, sizes, ins, dele):
    """One binary file with G_0 (CSR), the stream sizes [[n_add, n_del], ...] and
    the insert / delete sequences as (u, v) pairs."""
    row = np.ascontiguousarray(row, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64).reshape(-1)
    ins = np.ascontiguousarray(ins, dtype=np.uint32).reshape(-1)
    dele = np.ascontiguousarray(dele, dtype=np.uint32).reshape(-1)
    _check(lib().sg_write_snapshot_file(path.encode(), len(row) - 1, len(col), _ptr(row), _ptr(col), len(sizes) // 2,
                                        _ptr(sizes), _ptr(ins), _ptr(dele)))


def read_snapshot_file(path):
    """(row, col, sizes [nsnap, 2], ins [n, 2], dele [n, 2]) of a snapshot file."""
    nv, ne, ns = C.c_uint32(), C.c_uint64(), C.c_uint32()
    pr, pc, psz, pi, pd = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    _check(lib().sg_read_snapshot_file(path.encode(), C.byref(nv), C.byref(ne), C.byref(pr), C.byref(pc), C.byref(ns),
                                       C.byref(psz), C.byref(pi), C.byref(pd)))
    row = _take(pr, nv.value + 1, np.uint64)
    col = _take(pc, ne.value, np.uint32)
    sizes = _take(psz, 2 * ns.value, np.uint64).reshape(-1, 2)
    ti, td = int(sizes[:, 0].sum()), int(sizes[:, 1].sum())
    return row, col, sizes, _take(pi, 2 * ti, np.uint32).reshape(-1, 2), _take(pd, 2 * td, np.uint32).reshape(-1, 2)


def gen_rmat_device(scale, edge_factor, seed=DEFAULT_SEED, a=0.57, b=0.19, c=0.19, permute=False, device=0):
    """gen_rmat's graph, generated on a GPU: (nv, row, col)."""
    nv, ne, row, col = C.c_uint32(), C.c_uint64(), C.c_void_p(), C.c_void_p()
    _check(lib().sg_gen_rmat_device(device, scale, edge_factor, seed, a, b, c, int(permute), C.byref(nv), C.byref(ne),
                                    C.byref(row), C.byref(col)))
    return nv.value, _take(row, nv.value + 1, np.uint64), _take(col, ne.value, np.uint32)


def gen_rmat(scale, edge_factor, seed=DEFAULT_SEED, a=0.57, b=0.19, c=0.19, permute=False):
    nv, ne, row, col = C.c_uint32(), C.c_uint64(), C.c_void_p(), C.c_void_p()
    _check(lib().sg_gen_rmat(scale, edge_factor, seed, a, b, c, int(permute), C.byref(nv), C.byref(ne), C.byref(row), C.byref(col)))
    return nv.value, _take(row, nv.value + 1, np.uint64), _take(col, ne.value, np.uint32)


def gen_updates(row, col, scale, nsnap, n_ins, n_del, seed=DEFAULT_SEED, hub_frac=0.0, hub_top=0.01, a=0.57, b=0.19, c=0.19):
    row = np.ascontiguousarray(row, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    pi, pd = C.c_void_p(), C.c_void_p()
    _check(lib().sg_gen_updates(len(row) - 1, _ptr(row), _ptr(col), scale, a, b, c, nsnap, n_ins, n_del,
                                hub_frac, hub_top, seed, C.byref(pi), C.byref(pd)))
    ins = _take(pi, 2 * nsnap * n_ins, np.uint32).reshape(nsnap, n_ins, 2)
    dele = _take(pd, 2 * nsnap * n_del, np.uint32).reshape(nsnap, n_del, 2)
    return ins, dele


def edge_weight(mode, seed, u, v):
    return lib().sg_edge_weight(WEIGHT[mode] if isinstance(mode, str) else mode, seed, u, v)


# ---- engine -----------------------------------------------------------------
def default_config(**kw):
    cfg = Config()
    lib().sg_config_default(C.byref(cfg))
    for k, v in kw.items():
        if k == "algo" and isinstance(v, str):
            v = ALGO[v]
        if k == "weight_mode" and isinstance(v, str):
            v = WEIGHT[v]
        setattr(cfg, k, v)
    return cfg


class Engine:
    """One engine handle (single-threaded)."""

    def __init__(self, algo, create=None, **cfg):
        """algo: a built-in program ("pr", "sssp", ...), or a name for an engine
        made by `create` (an operator library's create function, load_operator)."""
        self.algo = algo
        self.cfg = default_config(algo=algo if create is None else "custom", **cfg)
        h = C.c_void_p()
        _check((create or lib().sg_create)(C.byref(self.cfg), C.byref(h)))
        self.h = h
        self.nv = 0

    def close(self):
        if getattr(self, "h", None):
            lib().sg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load_csr(self, row, col):
        row = np.ascontiguousarray(row, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        self.nv = len(row) - 1
        _check(lib().sg_load_csr(self.h, self.nv, len(col), _ptr(row), _ptr(col)), self.h)

    def init(self, source=None):
        if source is None:
            _check(lib().sg_init(self.h, None, 0), self.h)
        else:
            s = np.array([source], dtype=np.uint32)
            _check(lib().sg_init(self.h, _ptr(s), 1), self.h)

    def save_state(self, path):
        """Checkpoint: graph, vertex state and source to a file."""
        _check(lib().sg_save_state(self.h, path.encode()), self.h)

    def restore_state(self, path):
        """Load a checkpoint into this fresh engine (instead of load_csr + init)."""
        _check(lib().sg_restore_state(self.h, path.encode()), self.h)
        self.nv = int(self.stats()["nv"])

    def init_queries(self, sources, kinds):
        """Batched queries (engine created with algo "multi"): kinds are "sssp" / "bfs"."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        k = np.array([ALGO[x] if isinstance(x, str) else int(x) for x in kinds], dtype=np.uint32)
        if len(k) != len(src):
            raise ValueError("one kind per source")
        _check(lib().sg_init_queries(self.h, _ptr(src), _ptr(k), len(src)), self.h)

    def query_open(self, slots):
        """Continuous batching: fix `slots` query slots (drops every query)."""
        _check(lib().sg_query_open(self.h, slots), self.h)

    def query_add(self, sources, kinds):
        """Add queries into free slots; returns their slot numbers."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        k = np.array([ALGO[x] if isinstance(x, str) else int(x) for x in kinds], dtype=np.uint32)
        slots = np.empty(len(src), dtype=np.uint32)
        _check(lib().sg_query_add(self.h, _ptr(src), _ptr(k), len(src), _ptr(slots)), self.h)
        return slots

    def query_step(self, max_rounds, cap=4096):
        """Run up to max_rounds rounds: (slots that converged, slots still running)."""
        done = np.empty(cap, dtype=np.uint32)
        nd, na = C.c_uint32(), C.c_uint32()
        _check(lib().sg_query_step(self.h, max_rounds, _ptr(done), cap, C.byref(nd), C.byref(na)), self.h)
        return done[:nd.value].copy(), na.value

    def query_release(self, slots):
        s = np.ascontiguousarray(slots, dtype=np.uint32)
        _check(lib().sg_query_release(self.h, _ptr(s), len(s)), self.h)

    def query(self, qi, q=None):
        """Distances (or parents with q=Q_PARENT) of query qi of a batched engine; for
        "prcc" engines qi 0 = PageRank ranks (f32), qi 1 = CC labels."""
        out = np.empty(self.nv, dtype=np.float32 if (self.algo == "prcc" and qi == 0) else np.uint32)
        _check(lib().sg_gather_query(self.h, qi, Q_VALUE if q is None else q, _ptr(out), out.nbytes), self.h)
        return out

    def gather_at(self, vertices, q=None, qi=0):
        v = np.ascontiguousarray(vertices, dtype=np.uint32)
        out = np.empty(len(v), dtype=np.uint32)
        _check(lib().sg_gather_at(self.h, Q_VALUE if q is None else q, qi, _ptr(v), len(v), _ptr(out)), self.h)
        return out

    def gather_at_noalloc(self, vertices, q=None, qi=0):
        """gather_at through the library's fixed device scratch (no device allocation; <= 65536 ids)."""
        v = np.ascontiguousarray(vertices, dtype=np.uint32)
        out = np.empty(len(v), dtype=np.uint32)
        _check(lib().sg_gather_at_noalloc(self.h, Q_VALUE if q is None else q, qi, _ptr(v), len(v), _ptr(out)), self.h)
        return out

    def converge(self):
        """Converge snapshots deferred by FLAG_DEFER_CONVERGE (PageRank)."""
        _check(lib().sg_converge(self.h), self.h)

    def topk(self, k):
        ids = np.empty(k, dtype=np.uint32)
        ranks = np.empty(k, dtype=np.float32)
        _check(lib().sg_pr_topk(self.h, k, _ptr(ids), _ptr(ranks)), self.h)
        return ids, ranks

    def apply(self, ins, dele):
        ins = np.ascontiguousarray(ins, dtype=np.uint32).reshape(-1)
        dele = np.ascontiguousarray(dele, dtype=np.uint32).reshape(-1)
        _check(lib().sg_apply_snapshot(self.h, _ptr(ins), len(ins) // 2, _ptr(dele), len(dele) // 2), self.h)

    def apply_device(self, d_ins, nins, d_del, ndel):
        _check(lib().sg_apply_snapshot_device(self.h, C.c_void_p(d_ins), nins, C.c_void_p(d_del), ndel), self.h)

    def device_alloc(self, nbytes):
        p = C.c_void_p()
        _check(lib().sg_device_alloc(self.h, nbytes, C.byref(p)), self.h)
        return p.value

    def device_free(self, p):
        _check(lib().sg_device_free(self.h, C.c_void_p(p)), self.h)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.device_alloc(max(arr.nbytes, 1))
        if arr.nbytes:
            _check(lib().sg_copy_to_device(self.h, C.c_void_p(p), _ptr(arr), arr.nbytes), self.h)
        return p

    def gather(self, q):
        if q in (Q_OFFSETS, Q_IN_OFFSETS):
            out = np.empty(self.nv + 1, dtype=np.uint64)
        elif q == Q_SLOTS:
            out = np.empty(self.stats()["capacity"], dtype=np.uint32)
        elif q == Q_IN_SLOTS:
            out = np.empty(self.stats()["in_capacity"], dtype=np.uint32)
        elif q == Q_SEGMENTS:
            out = np.empty(max(self.cfg.segments, 1) + 1, dtype=np.uint32)
        elif q in (Q_VALUE, Q_BUFFER) and self.algo in ("pr", "prcc"):
            out = np.empty(self.nv, dtype=np.float32)
        else:
            out = np.empty(self.nv, dtype=np.uint32)
        _check(lib().sg_gather(self.h, q, _ptr(out), out.nbytes), self.h)
        return out

    def values(self):
        return self.gather(Q_VALUE)

    def adjacency(self):
        """Per-vertex adjacency sequences as a CSR (row, col) built from the PMA slots."""
        off = self.gather(Q_OFFSETS)
        deg = self.gather(Q_DEGREE)
        slots = self.gather(Q_SLOTS)
        row = np.zeros(self.nv + 1, dtype=np.uint64)
        np.cumsum(deg.astype(np.uint64), out=row[1:])
        idx = np.repeat(off[:-1], deg.astype(np.int64)) + (np.arange(int(row[-1]), dtype=np.uint64) - np.repeat(row[:-1], deg.astype(np.int64)))
        return row, slots[idx.astype(np.int64)], off, deg, slots

    def in_adjacency(self):
        """In-lists of the in-edge store as a CSR (row, col): the sources of each vertex's in-edges."""
        off = self.gather(Q_IN_OFFSETS)
        deg = self.gather(Q_IN_DEGREE)
        slots = self.gather(Q_IN_SLOTS)
        row = np.zeros(self.nv + 1, dtype=np.uint64)
        np.cumsum(deg.astype(np.uint64), out=row[1:])
        d = deg.astype(np.int64)
        idx = np.repeat(off[:-1].astype(np.int64), d) + (np.arange(int(row[-1]), dtype=np.int64) - np.repeat(row[:-1].astype(np.int64), d))
        return row, slots[idx], off, deg, slots

    def stats(self):
        s = Stats()
        _check(lib().sg_get_stats(self.h, C.byref(s)), self.h)
        return s.as_dict()

    def kernel_times(self):
        """{class: (launches, ms, algorithmic bytes)} of the HIP-event-timed kernel classes."""
        out = {}
        for c, name in enumerate(KCLASS):
            n, ms, b = C.c_uint64(), C.c_double(), C.c_uint64()
            _check(lib().sg_kernel_times(self.h, c, C.byref(n), C.byref(ms), C.byref(b)), self.h)
            out[name] = (n.value, ms.value, b.value)
        return out

    def reset_stats(self):
        _check(lib().sg_reset_stats(self.h), self.h)

    def synchronize(self):
        _check(lib().sg_synchronize(self.h), self.h)


class Server:
    """Request loop over one graph (sgraph_serve.h): updates, BFS/SSSP queries,
    CC label lookups and PageRank top-k, from any thread."""

    def __init__(self, row, col, query_slots=64, standing=SERVE_CC | SERVE_PR, hbm_reserve_gb=1.0, **engine):
        cfg = ServerConfig()
        lib().sg_server_config_default(C.byref(cfg))
        for k, v in engine.items():
            if k == "weight_mode" and isinstance(v, str):
                v = WEIGHT[v]
            setattr(cfg.engine, k, v)
        cfg.query_slots = query_slots
        cfg.standing = standing
        cfg.hbm_reserve_gb = hbm_reserve_gb
        row = np.ascontiguousarray(row, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        self.nv = len(row) - 1
        h = _P()
        st = lib().sg_server_create(C.byref(cfg), self.nv, len(col), _ptr(row), _ptr(col), C.byref(h))
        if st != SG_OK:
            raise SgError(st, lib().sg_server_last_error(None).decode())
        self.h = h

    def _err(self, st):
        if st != SG_OK:
            raise SgError(st, lib().sg_server_last_error(self.h).decode())

    def submit(self, kind, arg=0, target=INF, ins=None, dele=None):
        t = C.c_uint64()
        if kind == REQ_UPDATE:
            ins = np.ascontiguousarray(ins if ins is not None else [], dtype=np.uint32).reshape(-1)
            dele = np.ascontiguousarray(dele if dele is not None else [], dtype=np.uint32).reshape(-1)
            self._err(lib().sg_server_submit(self.h, kind, 0, 0, _ptr(ins), len(ins) // 2, _ptr(dele), len(dele) // 2,
                                             C.byref(t)))
        else:
            self._err(lib().sg_server_submit(self.h, kind, arg, target, None, 0, None, 0, C.byref(t)))
        return t.value

    def wait(self, ticket, kind=None, arg=0, target=INF):
        """Result array (or None for updates) and the ResultInfo as a dict."""
        info = ResultInfo()
        if kind in (REQ_BFS, REQ_SSSP):
            out = np.empty(self.nv if target == INF else 1, dtype=np.uint32)
        elif kind == REQ_CC:
            out = np.empty(1, dtype=np.uint32)
        elif kind == REQ_PR_TOPK:
            out = np.empty(2 * arg, dtype=np.uint32)
        else:
            out = None
        self._err(lib().sg_server_wait(self.h, ticket, _ptr(out) if out is not None else None,
                                       out.nbytes if out is not None else 0, C.byref(info)))
        if kind == REQ_PR_TOPK:
            out = (out[:arg].copy(), out[arg:].view(np.float32).copy())
        return out, {"snapshot": info.snapshot, "latency_ms": info.latency_ms, "step": info.step}

    def stats(self):
        s = ServerStats()
        self._err(lib().sg_server_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def close(self):
        if self.h:
            lib().sg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
