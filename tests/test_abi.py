"""The C-ABI library loads without a GPU and exports every entry point that
include/perigrain.h declares (no compute calls here)."""
import ctypes
import os
import re

from perigrain_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pg\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared("perigrain.h")
    for must in ("pg_create", "pg_batch_create", "pg_step", "pg_read_nodes", "pg_read_wall",
                 "pg_read_bond_intact", "pg_read_conn", "pg_read_contacts", "pg_destroy",
                 "pg_last_error"):
        assert must in names


def test_libpg_exports_every_declared_symbol():
    lib = ctypes.CDLL(_ffi.LIBPG_PATH)
    missing = [n for n in declared("perigrain.h") if not hasattr(lib, n)]
    assert not missing, f"libpg.so lacks {missing}"


def test_device_count_without_gpu_is_clean():
    # returns an integer (0 in a container without a GPU) instead of failing
    assert _ffi.device_count() >= 0


def test_create_rejects_bad_setup():
    import numpy as np
    import pytest
    from perigrain_amd import Stepper, setups
    s, p = setups.two_disk_collision(seed=0)
    s.conn = np.concatenate([s.conn, s.conn[:1]])  # a bond listed twice
    s.conn_offset = s.conn_offset.copy()
    s.conn_offset[-1] += 1
    with pytest.raises(_ffi.PgError, match="twice"):
        Stepper(s, p, device=0)
    s2, p2 = setups.two_disk_collision(seed=0)
    s2.conn[5, 1] = s2.node_offset[1] + 7  # index beyond grain 0
    with pytest.raises(_ffi.PgError, match="out of range"):
        Stepper(s2, p2, device=0)


def test_libpgio_exports_every_declared_symbol():
    from perigrain_amd import h5io
    lib = ctypes.CDLL(h5io.LIBPGIO_PATH)
    missing = [n for n in declared("pgio.h") if not hasattr(lib, n)]
    assert declared("pgio.h") and not missing, f"libpgio.so lacks {missing}"


def test_create_rejects_bad_loop_settings():
    import pytest
    from perigrain_amd import Stepper, setups
    s, p = setups.two_disk_collision(seed=0)
    p.timeloop_variant = 3
    with pytest.raises(_ffi.PgError, match="timeloop_variant"):
        Stepper(s, p, device=0)
    s, p = setups.two_disk_collision(seed=0)
    s = setups.subset(s, [0])
    p.reset_partzero_y, p.reset_partzero_y_timestep = 1, 10
    with pytest.raises(_ffi.PgError, match="second grain"):
        Stepper(s, p, device=0)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_null_handle_is_an_error_without_a_device():
    lib = _ffi.lib()
    invalid = 1  # PG_ERR_INVALID
    assert lib.pg_step(None, 1) == invalid
    assert b"null handle" in lib.pg_last_error()
    assert lib.pg_sync(None) == invalid


def test_adjoint_keeps_the_whole_run_when_it_fits():
    # diffrun's default checkpoint interval: the whole run (no recompute in
    # the backward pass) while every step's state fits the budget, else
    # checkpoints every ceil(sqrt(steps)) steps
    from perigrain_amd import diffrun, setups
    s, _ = setups.compression_2d_diff(n_grains=4, nodes_per_grain=60)
    per_step = diffrun.trajectory_bytes(s, 1)
    ns = (s.n_nodes + 1 + 63) // 64 * 64
    assert per_step == ns * 24 * s.dim + s.n_grains * 2 * s.dim * 8 + 48
    assert diffrun.auto_checkpoint_every(s, 2000) == 2000
    assert diffrun.auto_checkpoint_every(s, 2000, budget=per_step * 1999) == 45
    assert diffrun.auto_checkpoint_every(s, 2000, budget=per_step * 2000) == 2000
