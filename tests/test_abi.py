"""The C ABI library loads and exports every symbol include/*.h declares;
argument and state errors come back as status codes (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import sgraph
from conftest import ROOT


def declared_functions():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(sg_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = sgraph.lib()
    names = declared_functions()
    assert len(names) >= 36
    for n in names:
        assert hasattr(lib, n), n
    # and the binding covers them all
    assert set(names) == {s[0] for s in sgraph.SIGNATURES}


def test_config_defaults_are_the_references():
    cfg = sgraph.default_config()
    assert cfg.pr_alpha == 0.85 and cfg.pr_epsilon == 1e-6
    assert cfg.pma_density == 0.25 and cfg.n_stream == 3 and cfg.pull_divisor == 0
    assert cfg.segments == 1 and cfg.weight_seed == 27491095
    assert cfg.cache_level == 4 and cfg.edge_budget_gb == 0  # app_skeleton.h:100 -cache 4


def test_struct_layouts_match_the_header():
    # the ctypes mirrors must have the C structs' sizes (sg_config 104 bytes, sg_stats 240)
    assert C.sizeof(sgraph.Config) == 104
    assert C.sizeof(sgraph.Stats) == 240


def test_server_config_defaults():
    cfg = sgraph.ServerConfig()
    sgraph.lib().sg_server_config_default(C.byref(cfg))
    assert cfg.query_slots == 64 and cfg.standing == 3 and cfg.hbm_reserve_gb == 1.0
    assert cfg.engine.pr_alpha == 0.85


def test_null_handles_are_einval():
    lib = sgraph.lib()
    assert lib.sg_init(None, None, 0) == sgraph.SG_EINVAL
    assert lib.sg_apply_snapshot(None, None, 0, None, 0) == sgraph.SG_EINVAL
    assert lib.sg_create(None, None) == sgraph.SG_EINVAL
    assert lib.sg_last_error(None)
    lib.sg_destroy(None)


def test_bad_config_is_einval():
    cfg = sgraph.default_config(algo=9)
    h = C.c_void_p()
    assert sgraph.lib().sg_create(C.byref(cfg), C.byref(h)) == sgraph.SG_EINVAL
    cfg = sgraph.default_config(algo=0, pr_alpha=1.5)
    assert sgraph.lib().sg_create(C.byref(cfg), C.byref(h)) == sgraph.SG_EINVAL


def test_create_without_device_fails_cleanly():
    cfg = sgraph.default_config(algo=1, device=99)
    h = C.c_void_p()
    st = sgraph.lib().sg_create(C.byref(cfg), C.byref(h))
    assert st == sgraph.SG_EDEVICE
    assert b"device" in sgraph.lib().sg_last_error(None).lower() or sgraph.lib().sg_last_error(None)


def test_server_create_failure_reports_its_reason():
    # device 99 does not exist (here no device does): the server's engines
    # cannot be made, and the reason is the calling thread's last server error
    nv, row, col = sgraph.gen_rmat(6, 4, seed=1)
    cfg = sgraph.ServerConfig()
    sgraph.lib().sg_server_config_default(C.byref(cfg))
    cfg.engine.device = 99
    h = C.c_void_p()
    row = np.ascontiguousarray(row, dtype=np.uint64)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    st = sgraph.lib().sg_server_create(C.byref(cfg), nv, len(col), sgraph._ptr(row), sgraph._ptr(col), C.byref(h))
    assert st == sgraph.SG_EDEVICE and not h.value
    msg = sgraph.lib().sg_server_last_error(None)
    assert msg and b"device" in msg.lower()


def test_device_generator_without_device_fails_cleanly():
    with pytest.raises(sgraph.SgError) as e:
        sgraph.gen_rmat_device(6, 4, device=99)
    assert e.value.status == sgraph.SG_EDEVICE


def test_operator_library_loads_and_refuses_without_device():
    # the widest-path operator library (include/sgraph_operator.h) exports its
    # create function; without a device the create fails with a status and
    # libsgraph's sg_last_error(NULL) explains it
    f = sgraph.load_operator("widest_path")
    cfg = sgraph.default_config(algo="custom", device=99)
    h = C.c_void_p()
    assert f(C.byref(cfg), C.byref(h)) == sgraph.SG_EDEVICE and not h.value
    assert b"device" in sgraph.lib().sg_last_error(None).lower()
    cfg = sgraph.default_config(algo="custom", weight_mode=7)
    assert f(C.byref(cfg), C.byref(h)) == sgraph.SG_EINVAL
    assert f(None, C.byref(h)) == sgraph.SG_EINVAL


def test_sg_create_refuses_custom_algo():
    cfg = sgraph.default_config(algo="custom")
    h = C.c_void_p()
    assert sgraph.lib().sg_create(C.byref(cfg), C.byref(h)) == sgraph.SG_EINVAL
    assert b"operator library" in sgraph.lib().sg_last_error(None)
