"""Parity oracle: ctypes binding of liboracle.so (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY.  The oracle is a plain FP64 CPU restatement of the
reference time loop (src/particle/timeloop.h:500-1015 and the functions it
calls).  It is imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the CPU baseline -- never by
the perigrain_amd package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force=False):
    """Compile liboracle.so with the reference's arithmetic flags."""
    if os.path.exists(LIB_PATH) and not force:
        return LIB_PATH
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    from perigrain_amd import _ffi
    H = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.oracle_last_error.restype = ctypes.c_char_p
    sig = {
        "oracle_create": [ctypes.POINTER(_ffi.PgSetup), ctypes.POINTER(_ffi.PgParams),
                          ctypes.POINTER(H)],
        "oracle_destroy": [H],
        "oracle_step": [H, ctypes.c_uint32, ctypes.c_int, ctypes.c_int],
        "oracle_sizes": [H, u64p, u64p, u64p],
        "oracle_read_nodes": [H, ctypes.c_int, dp],
        "oracle_read_wall": [H, dp, dp],
        "oracle_read_bond_intact": [H, ctypes.POINTER(ctypes.c_uint8)],
        "oracle_read_conn": [H, ctypes.c_int, u32p, ctypes.c_uint64, u64p],
        "oracle_read_contacts": [H, u32p, ctypes.c_uint64, u64p],
        "oracle_max_threads": [],
        "oracle_create_tangent": [ctypes.POINTER(_ffi.PgSetup), ctypes.POINTER(_ffi.PgParams),
                                  ctypes.POINTER(H)],
        "oracle_read_wall_tangent": [H, dp, dp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.oracle_segment_mass.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp]
    L.oracle_segment_mass.restype = None
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().oracle_last_error().decode()}")


class Oracle:
    """The reference time loop on the CPU, same interface as Stepper."""

    def __init__(self, setup, params, threads=1, tangent=False):
        from perigrain_amd import _ffi
        L = lib()
        st, self._keep = _ffi.setup_struct(setup)
        pp = _ffi.params_struct(params)
        h = ctypes.c_void_p()
        create = L.oracle_create_tangent if tangent else L.oracle_create
        _check(create(ctypes.byref(st), ctypes.byref(pp), ctypes.byref(h)))
        self._h = h
        self.dim = int(setup.dim)
        self.threads = int(threads)
        self.t = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1, record_contacts=False):
        _check(lib().oracle_step(self._h, int(n), self.threads, 1 if record_contacts else 0))
        self.t += int(n)

    def sizes(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().oracle_sizes(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def read(self, field):
        from perigrain_amd import _ffi
        n, _, _ = self.sizes()
        out = np.empty((n, self.dim))
        _check(lib().oracle_read_nodes(self._h, _ffi.FIELDS[field],
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def wall(self):
        info = np.zeros(6)
        react = np.zeros(6 * self.dim)
        _check(lib().oracle_read_wall(self._h, info.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      react.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        ns = 4 if self.dim == 2 else 6
        return info[:ns].copy(), react.reshape(6, self.dim)[:ns].copy()

    def wall_tangent(self):
        """Reaction of the last step (sides, dim) and, for a tangent-mode
        oracle, its derivatives (sides, dim, 3) along (cnot scale,
        normal_stiffness, friction_coefficient)."""
        ns = 4 if self.dim == 2 else 6
        r = np.zeros(ns * self.dim)
        dr = np.zeros(ns * self.dim * 3)
        _check(lib().oracle_read_wall_tangent(self._h, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                              dr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return r.reshape(ns, self.dim), dr.reshape(ns, self.dim, 3)

    def bond_intact(self):
        _, _, ns = self.sizes()
        out = np.empty(ns, np.uint8)
        _check(lib().oracle_read_bond_intact(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def conn(self, grain):
        rows = ctypes.c_uint64()
        _check(lib().oracle_read_conn(self._h, grain, None, 0, ctypes.byref(rows)))
        out = np.empty((rows.value, 2), np.uint32)
        _check(lib().oracle_read_conn(self._h, grain, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                      rows.value, ctypes.byref(rows)))
        return out

    def contacts(self):
        n = ctypes.c_uint64()
        _check(lib().oracle_read_contacts(self._h, None, 0, ctypes.byref(n)))
        out = np.empty((n.value, 2), np.uint32)
        if n.value:
            _check(lib().oracle_read_contacts(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                              n.value, ctypes.byref(n)))
        return out


def reaction_history_tangent(setup, params, steps, threads=1):
    """FP64 CPU reference (dual numbers): the wall reaction of every step,
    (steps, sides, dim), and its derivatives along (cnot scale,
    normal_stiffness, friction_coefficient), (steps, sides, dim, 3)."""
    o = Oracle(setup, params, threads=threads, tangent=True)
    try:
        ns = 4 if setup.dim == 2 else 6
        R = np.zeros((steps, ns, setup.dim))
        dR = np.zeros((steps, ns, setup.dim, 3))
        for t in range(steps):
            o.step(1)
            R[t], dR[t] = o.wall_tangent()
        return R, dR
    finally:
        o.close()


def segment_mass(dim, R, d):
    c, a = ctypes.c_double(), ctypes.c_double()
    lib().oracle_segment_mass(int(dim), float(R), float(d), ctypes.byref(c), ctypes.byref(a))
    return c.value, a.value
