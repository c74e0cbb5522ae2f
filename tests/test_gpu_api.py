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


def test_snapshot_slots_read_while_stepping():
    # a snapshot slot holds exactly what the plain readers gave right after
    # its step, read back on another thread while the handle steps on
    import threading
    s, p = setups.two_disk_collision(seed=4, gap=1e-5, speed=50.0)
    nodes = s.grain_nodes(0)
    s.vel[nodes] += 1e5 * (s.pos[nodes] - s.pos[nodes].mean(axis=0))  # bonds break
    fields = ("CurrPos", "vel", "acc", "force", "disp")
    with Stepper(s, p, device=0) as g, Stepper(s, p, device=0) as ref:
        want = {}
        for slot, t in ((0, 40), (1, 90)):
            g.step(t - g.t)
            ref.step(t - ref.t)
            g.snapshot_take(slot)
            want[slot] = ({f: ref.read(f) for f in fields},
                          [ref.conn(k) for k in range(s.n_grains)], ref.wall())
        got = {}

        def reader():
            for slot in (0, 1):
                got[slot] = ({f: g.snapshot_read(f, slot) for f in fields},
                             [g.snapshot_conn(k, slot) for k in range(s.n_grains)],
                             g.snapshot_wall(slot), g.snapshot_step(slot))

        th = threading.Thread(target=reader)
        th.start()
        g.step(200)  # the live state moves on meanwhile
        th.join()
        assert g.t == 290
        for slot, t in ((0, 40), (1, 90)):
            wf, wc, (ww, wr) = want[slot]
            gf, gc, (gw, gr), gt = got[slot]
            assert gt == t
            for f in fields:
                assert np.array_equal(gf[f], wf[f]), (slot, f)
            for k in range(s.n_grains):
                assert np.array_equal(gc[k], wc[k]), (slot, k)
            assert np.array_equal(gw, ww) and np.array_equal(gr, wr)
        assert len(want[1][1][0]) < len(want[0][1][0]), "bonds should break between the slots"
        with pytest.raises(_ffi.PgError):
            g.snapshot_read("vel", 2)
