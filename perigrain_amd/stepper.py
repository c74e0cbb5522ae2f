"""Python driver of libpg: the HIP time stepper behind the C ABI.

`Stepper` mirrors run_timeloop's role (src/particle/timeloop.h:500-1015):
it owns the packed device state of one simulation (or a batch) and advances
it step by step; snapshots are read back in the reference's layouts.
"""
import atexit
import ctypes
import weakref

import numpy as np

from . import _ffi
from ._ffi import check, lib

# Handles still open when the interpreter exits are destroyed by this hook,
# which runs before the interpreter tears down its modules and before the
# HIP runtime's own exit handlers, rather than by __del__ during or after
# that teardown.
_live = weakref.WeakSet()


@atexit.register
def _close_live_handles():
    for st in list(_live):
        st.close()


class Stepper:
    """One simulation (or a batch of independent ones) on a HIP device."""

    def __init__(self, setups, params, device=0):
        single = not isinstance(setups, (list, tuple))
        if single:
            setups, params = [setups], [params]
        if len(setups) != len(params):
            raise ValueError("setups and params must have the same length")
        L = lib()
        self._structs = []
        n = len(setups)
        sarr = (_ffi.PgSetup * n)()
        parr = (_ffi.PgParams * n)()
        for k, (s, p) in enumerate(zip(setups, params)):
            st, keep = _ffi.setup_struct(s)
            self._structs.append(keep)
            sarr[k] = st
            parr[k] = _ffi.params_struct(p)
        h = ctypes.c_void_p()
        check(L.pg_batch_create(sarr, parr, n, int(device), ctypes.byref(h)))
        self._h = h
        _live.add(self)
        self.dim = int(setups[0].dim)
        self.n_instances = n
        self.setups = setups
        self.params = params

    # ------------------------------------------------------------- control
    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            _live.discard(self)
            lib().pg_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n=1):
        check(lib().pg_step(self._h, int(n)))

    def sync(self):
        check(lib().pg_sync(self._h))

    @property
    def t(self):
        v = ctypes.c_uint32()
        check(lib().pg_current_step(self._h, ctypes.byref(v)))
        return int(v.value)

    def sizes(self, inst=0):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().pg_sizes(self._h, inst, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    # ------------------------------------------------------------ readback
    def read(self, field, inst=0):
        n, _, _ = self.sizes(inst)
        out = np.empty((n, self.dim), dtype=np.float64)
        check(lib().pg_read_nodes(self._h, inst, _ffi.FIELDS[field],
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def wall(self, inst=0):
        info = np.zeros(6)
        react = np.zeros(6 * self.dim)
        check(lib().pg_read_wall(self._h, inst, info.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 react.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        ns = 4 if self.dim == 2 else 6
        return info[:ns].copy(), react.reshape(6, self.dim)[:ns].copy()

    def bond_intact(self, inst=0):
        _, _, ns = self.sizes(inst)
        out = np.empty(ns, dtype=np.uint8)
        check(lib().pg_read_bond_intact(self._h, inst,
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def grain_stats(self, inst=0):
        """Per-grain post-processing of the last completed step on the device
        (pg_grain_stats; read_sim_output.py extract_bulk / read_force_val):
        arrays over the instance's grains in reference order."""
        _, ng, _ = self.sizes(inst)
        d = self.dim
        out = np.empty((ng, 5 * d + 1), dtype=np.float64)
        check(lib().pg_grain_stats(self._h, inst,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return {"force_sum": out[:, :d], "mean_vel": out[:, d:2 * d],
                "mean_acc": out[:, 2 * d:3 * d], "mean_CurrPos": out[:, 3 * d:4 * d],
                "damage": out[:, 4 * d], "vol_rho_acc_sum": out[:, 4 * d + 1:]}

    def conn(self, grain, inst=0):
        rows = ctypes.c_uint64()
        check(lib().pg_read_conn(self._h, inst, grain, None, 0, ctypes.byref(rows)))
        out = np.empty((rows.value, 2), dtype=np.uint32)
        check(lib().pg_read_conn(self._h, inst, grain,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                 rows.value, ctypes.byref(rows)))
        return out

    def contacts(self, inst=0):
        n = ctypes.c_uint64()
        check(lib().pg_read_contacts(self._h, inst, None, 0, ctypes.byref(n)))
        out = np.empty((n.value, 2), dtype=np.uint32)
        if n.value:
            check(lib().pg_read_contacts(self._h, inst,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                         n.value, ctypes.byref(n)))
        return out

    def stats(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().pg_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"rebuilds": a.value, "list_entries": b.value, "skipped_steps": c.value}

    def set_tuning(self, skin_factor=0.5, block=256):
        check(lib().pg_set_tuning(self._h, float(skin_factor), int(block)))

    def set_graph(self, on=True):
        check(lib().pg_set_graph(self._h, 1 if on else 0))

    # -------------------------------------------------------------- timing
    def event(self, slot):
        check(lib().pg_event_record(self._h, slot))

    def elapsed_ms(self, a, b):
        ms = ctypes.c_float()
        check(lib().pg_event_elapsed(self._h, a, b, ctypes.byref(ms)))
        return float(ms.value)

    def autotune(self, trial_steps=64):
        """Pick the fastest (skin factor, block) by running 9*trial_steps
        real steps (pg_autotune); returns the choice."""
        sk = ctypes.c_double()
        bl = ctypes.c_int()
        check(lib().pg_autotune(self._h, int(trial_steps), ctypes.byref(sk), ctypes.byref(bl)))
        return float(sk.value), int(bl.value)

    def adjoint_forward(self, nsteps, checkpoint_every=100):
        """Run nsteps recorded steps (pg_adjoint_forward); returns the wall
        reaction of every step, shape (nsteps, sides, dim)."""
        ns = 4 if self.dim == 2 else 6
        out = np.zeros((int(nsteps), 6, self.dim))
        check(lib().pg_adjoint_forward(self._h, int(nsteps), int(checkpoint_every),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out[:, :ns].copy()

    def adjoint_backward(self, reaction_bar):
        """Cotangent of the recorded reaction history -> derivatives with
        respect to (micromodulus scale, normal stiffness, friction)."""
        ns = 4 if self.dim == 2 else 6
        rb = np.asarray(reaction_bar, dtype=np.float64)
        full = np.zeros((rb.shape[0], 6, self.dim))
        full[:, :ns] = rb
        g = np.zeros(3)
        check(lib().pg_adjoint_backward(self._h, full.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return g

    # ------------------------------------------------------------ snapshots
    def snapshot_take(self, slot=0):
        """Copy the last completed step's snapshot state into slot `slot` on
        the device, in stream order (pg_snapshot_take); returns at once."""
        check(lib().pg_snapshot_take(self._h, int(slot)))

    def snapshot_step(self, slot=0):
        v = ctypes.c_uint32()
        check(lib().pg_snapshot_step(self._h, int(slot), ctypes.byref(v)))
        return int(v.value)

    def snapshot_read(self, field, slot=0, inst=0):
        """A node field of a taken snapshot slot; may be called from another
        thread while this handle steps."""
        n, _, _ = self.sizes(inst)
        out = np.empty((n, self.dim), dtype=np.float64)
        check(lib().pg_snapshot_read_nodes(self._h, int(slot), inst, _ffi.FIELDS[field],
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def snapshot_conn(self, grain, slot=0, inst=0):
        rows = ctypes.c_uint64()
        check(lib().pg_snapshot_read_conn(self._h, int(slot), inst, grain, None, 0,
                                          ctypes.byref(rows)))
        out = np.empty((rows.value, 2), dtype=np.uint32)
        check(lib().pg_snapshot_read_conn(self._h, int(slot), inst, grain,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                          rows.value, ctypes.byref(rows)))
        return out

    def snapshot_wall(self, slot=0, inst=0):
        info = np.zeros(6)
        react = np.zeros(6 * self.dim)
        check(lib().pg_snapshot_read_wall(self._h, int(slot), inst,
                                          info.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          react.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        ns = 4 if self.dim == 2 else 6
        return info[:ns].copy(), react.reshape(6, self.dim)[:ns].copy()

    def kernel_time(self):
        """(milliseconds, launches) of step-kernel runs so far (pg_kernel_time)."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        check(lib().pg_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return float(ms.value), int(n.value)
