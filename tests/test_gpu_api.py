"""Edge cases and error paths of the C ABI on a GPU: the state before any
step, steps split over several calls, argument errors, the capacity rule of
the readers and the step-kernel timer."""
import ctypes

import numpy as np
import pytest

from perigrain_amd import Stepper, _ffi, setups

pytestmark = pytest.mark.gpu


def test_state_before_any_step():
    s, p = setups.two_disk_collision(seed=2, gap=5e-5)
    with Stepper(s, p, device=0) as g:
        g.step(0)
        assert g.t == 0
        assert np.array_equal(g.read("CurrPos"), s.pos + s.disp)
        assert np.array_equal(g.read("disp"), s.disp)
        assert np.array_equal(g.read("vel"), s.vel)
        assert np.array_equal(g.read("acc"), s.acc)
        assert g.bond_intact().all()


def test_steps_split_across_calls_equal_one_call():
    # call boundaries fall inside and at the ends of the 64-step batches
    s, p = setups.two_disk_collision(seed=3, gap=5e-5, speed=20.0)
    with Stepper(s, p, device=0) as a, Stepper(s, p, device=0) as b:
        a.step(300)
        for n in (1, 36, 64, 65, 134):
            b.step(n)
        assert a.t == b.t == 300
        for f in ("CurrPos", "vel", "acc", "force"):
            assert np.array_equal(a.read(f), b.read(f)), f
        assert np.array_equal(a.contacts(), b.contacts())
        assert len(a.contacts()) > 0
        wa, ra = a.wall()
        wb, rb = b.wall()
        assert np.array_equal(wa, wb) and np.array_equal(ra, rb)


def test_bad_arguments_are_errors():
    s, p = setups.two_disk_collision(seed=2)
    with Stepper(s, p, device=0) as g:
        with pytest.raises(_ffi.PgError, match="instance"):
            g.read("vel", inst=1)
        buf = np.empty(s.n_nodes * 2)
        with pytest.raises(_ffi.PgError, match="unknown field"):
            _ffi.check(_ffi.lib().pg_read_nodes(
                g._h, 0, 9, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        with pytest.raises(_ffi.PgError, match="bad grain"):
            g.conn(5)
        with pytest.raises(_ffi.PgError, match="block"):
            g.set_tuning(0.5, 100)
        with pytest.raises(_ffi.PgError, match="skin_factor"):
            g.set_tuning(0.0, 256)
        # the handle still works after the errors
        g.step(5)
        assert g.t == 5


def test_conn_reader_truncates_to_the_given_capacity():
    s, p = setups.two_disk_collision(seed=2)
    with Stepper(s, p, device=0) as g:
        full = g.conn(0)
        rows = ctypes.c_uint64()
        part = np.empty((5, 2), dtype=np.uint32)
        _ffi.check(_ffi.lib().pg_read_conn(
            g._h, 0, 0, part.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 5,
            ctypes.byref(rows)))
        assert rows.value == len(full)
        assert np.array_equal(part, full[:5])


def test_kernel_timer_counts_step_launches():
    s, p = setups.two_disk_collision(seed=2)
    with Stepper(s, p, device=0) as g:
        g.step(130)
        ms, n = g.kernel_time()
        assert 0 < n <= 130 and ms > 0
