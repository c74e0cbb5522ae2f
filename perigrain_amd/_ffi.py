"""ctypes view of the libpg C ABI (include/perigrain.h).

libpg.so is built in-tree next to this file by ``perigrain_amd.build``.
Loading fails loudly when it is missing: there is no CPU fallback in the
product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PERIGRAIN_LIBPG selects another build of the library (e.g. a tuning variant)
LIBPG_PATH = os.environ.get("PERIGRAIN_LIBPG") or os.path.join(_HERE, "libpg.so")

PG_OK = 0
PG_MAX_SIDES = 6
FIELDS = {"CurrPos": 0, "vel": 1, "acc": 2, "force": 3, "disp": 4}

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class PgSetup(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("n_grains", ctypes.c_int32),
        ("node_offset", _i64p),
        ("pos", _dp),
        ("vol", _dp),
        ("disp", _dp),
        ("vel", _dp),
        ("acc", _dp),
        ("extforce", _dp),
        ("conn_offset", _i64p),
        ("conn", _u32p),
        ("bdry_offset", _i64p),
        ("bdry_nodes", _u32p),
        ("clamped_offset", _i64p),
        ("clamped_nodes", _u32p),
        ("delta", _dp),
        ("rho", _dp),
        ("cnot", _dp),
        ("snot", _dp),
        ("movable", _i32p),
        ("breakable", _i32p),
        ("stoppable", _i32p),
        ("torque_axis", _u32p),
        ("torque_val", _dp),
        ("contact_radius", ctypes.c_double),
        ("normal_stiffness", ctypes.c_double),
        ("damping_ratio", ctypes.c_double),
        ("friction_coefficient", ctypes.c_double),
        ("allow_wall", ctypes.c_int32),
        ("wall", ctypes.c_double * PG_MAX_SIDES),
    ]


class PgParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("enable_fracture", ctypes.c_int32),
        ("nl_bdry_only", ctypes.c_int32),
        ("allow_damping", ctypes.c_int32),
        ("allow_friction", ctypes.c_int32),
        ("self_contact", ctypes.c_int32),
        ("gradient_extforce", ctypes.c_int32),
        ("extforce_maxstep", ctypes.c_uint32),
        ("enable_torque", ctypes.c_int32),
        ("set_movable_index", ctypes.c_int32),
        ("set_movable_timestep", ctypes.c_int32),
        ("set_stoppable_index", ctypes.c_int32),
        ("set_stoppable_timestep", ctypes.c_int32),
        ("reset_partzero_y", ctypes.c_int32),
        ("reset_partzero_y_timestep", ctypes.c_uint32),
        ("wheel_rad", ctypes.c_double),
        ("wall_speed", ctypes.c_double * PG_MAX_SIDES),
        ("timeloop_variant", ctypes.c_int32),
    ]


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def setup_struct(s):
    """Build a PgSetup from a perigrain_amd.setups.Setup.

    Returns (struct, keepalive) -- keep the second item alive while the
    struct is in use.
    """
    keep = []

    def arr(x, dtype):
        y = np.ascontiguousarray(x, dtype=dtype)
        keep.append(y)
        return y

    st = PgSetup()
    st.dim = int(s.dim)
    st.n_grains = int(len(s.node_offset) - 1)
    st.node_offset = _ptr(arr(s.node_offset, np.int64), ctypes.c_int64)
    st.pos = _ptr(arr(s.pos, np.float64), ctypes.c_double)
    st.vol = _ptr(arr(s.vol, np.float64), ctypes.c_double)
    st.disp = _ptr(arr(s.disp, np.float64), ctypes.c_double)
    st.vel = _ptr(arr(s.vel, np.float64), ctypes.c_double)
    st.acc = _ptr(arr(s.acc, np.float64), ctypes.c_double)
    st.extforce = _ptr(arr(s.extforce, np.float64), ctypes.c_double)
    st.conn_offset = _ptr(arr(s.conn_offset, np.int64), ctypes.c_int64)
    st.conn = _ptr(arr(s.conn.reshape(-1) if len(s.conn) else np.zeros(2), np.uint32), ctypes.c_uint32)
    st.bdry_offset = _ptr(arr(s.bdry_offset, np.int64), ctypes.c_int64)
    st.bdry_nodes = _ptr(arr(s.bdry_nodes if len(s.bdry_nodes) else np.zeros(1), np.uint32), ctypes.c_uint32)
    st.clamped_offset = _ptr(arr(s.clamped_offset, np.int64), ctypes.c_int64)
    st.clamped_nodes = _ptr(arr(s.clamped_nodes if len(s.clamped_nodes) else np.zeros(1), np.uint32),
                            ctypes.c_uint32)
    st.delta = _ptr(arr(s.delta, np.float64), ctypes.c_double)
    st.rho = _ptr(arr(s.rho, np.float64), ctypes.c_double)
    st.cnot = _ptr(arr(s.cnot, np.float64), ctypes.c_double)
    st.snot = _ptr(arr(s.snot, np.float64), ctypes.c_double)
    st.movable = _ptr(arr(s.movable, np.int32), ctypes.c_int32)
    st.breakable = _ptr(arr(s.breakable, np.int32), ctypes.c_int32)
    st.stoppable = _ptr(arr(s.stoppable, np.int32), ctypes.c_int32)
    st.torque_axis = _ptr(arr(s.torque_axis, np.uint32), ctypes.c_uint32)
    st.torque_val = _ptr(arr(s.torque_val, np.float64), ctypes.c_double)
    st.contact_radius = float(s.contact_radius)
    st.normal_stiffness = float(s.normal_stiffness)
    st.damping_ratio = float(s.damping_ratio)
    st.friction_coefficient = float(s.friction_coefficient)
    st.allow_wall = int(s.allow_wall)
    for k in range(PG_MAX_SIDES):
        st.wall[k] = float(s.wall[k])
    return st, keep


def params_struct(p):
    pp = PgParams()
    pp.dt = float(p.dt)
    for name in ("enable_fracture", "nl_bdry_only", "allow_damping", "allow_friction",
                 "self_contact", "gradient_extforce", "enable_torque", "set_movable_index",
                 "set_movable_timestep", "set_stoppable_index", "set_stoppable_timestep",
                 "reset_partzero_y"):
        setattr(pp, name, int(getattr(p, name)))
    pp.extforce_maxstep = int(p.extforce_maxstep)
    pp.reset_partzero_y_timestep = int(p.reset_partzero_y_timestep)
    pp.wheel_rad = float(p.wheel_rad)
    for k in range(PG_MAX_SIDES):
        pp.wall_speed[k] = float(p.wall_speed[k])
    pp.timeloop_variant = int(getattr(p, "timeloop_variant", 0))
    return pp


_lib = None


def lib():
    """Load libpg.so (raises if the HIP extension has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPG_PATH):
        raise RuntimeError(
            f"libpg.so not found at {LIBPG_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or "
            "`python -m perigrain_amd.build`")
    L = ctypes.CDLL(LIBPG_PATH)
    H = ctypes.c_void_p
    L.pg_last_error.restype = ctypes.c_char_p
    L.pg_last_error.argtypes = []
    sig = {
        "pg_version": [ctypes.c_char_p, ctypes.c_size_t],
        "pg_create": [ctypes.POINTER(PgSetup), ctypes.POINTER(PgParams), ctypes.c_int,
                      ctypes.POINTER(H)],
        "pg_batch_create": [ctypes.POINTER(PgSetup), ctypes.POINTER(PgParams), ctypes.c_int,
                            ctypes.c_int, ctypes.POINTER(H)],
        "pg_destroy": [H],
        "pg_step": [H, ctypes.c_uint32],
        "pg_sync": [H],
        "pg_current_step": [H, _u32p],
        "pg_sizes": [H, ctypes.c_int, _u64p, _u64p, _u64p],
        "pg_read_nodes": [H, ctypes.c_int, ctypes.c_int, _dp],
        "pg_read_wall": [H, ctypes.c_int, _dp, _dp],
        "pg_read_bond_intact": [H, ctypes.c_int, _u8p],
        "pg_grain_stats": [H, ctypes.c_int, _dp],
        "pg_grain_stats_noalloc": [H, ctypes.c_int, _dp],
        "pg_read_conn": [H, ctypes.c_int, ctypes.c_int, _u32p, ctypes.c_uint64, _u64p],
        "pg_read_contacts": [H, ctypes.c_int, _u32p, ctypes.c_uint64, _u64p],
        "pg_stream": [H, ctypes.POINTER(ctypes.c_void_p)],
        "pg_stats": [H, _u64p, _u64p, _u64p],
        "pg_set_tuning": [H, ctypes.c_double, ctypes.c_int],
        "pg_set_graph": [H, ctypes.c_int],
        "pg_autotune": [H, ctypes.c_uint32, _dp, ctypes.POINTER(ctypes.c_int)],
        "pg_event_record": [H, ctypes.c_int],
        "pg_event_elapsed": [H, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)],
        "pg_kernel_time": [H, _dp, _u64p],
        "pg_snapshot_take": [H, ctypes.c_int],
        "pg_snapshot_step": [H, ctypes.c_int, _u32p],
        "pg_snapshot_read_nodes": [H, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp],
        "pg_snapshot_read_conn": [H, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p,
                                  ctypes.c_uint64, _u64p],
        "pg_snapshot_read_wall": [H, ctypes.c_int, ctypes.c_int, _dp, _dp],
        "pg_adjoint_forward": [H, ctypes.c_uint32, ctypes.c_uint32, _dp],
        "pg_adjoint_backward": [H, _dp, _dp],
        "pg_device_count": [ctypes.POINTER(ctypes.c_int)],
        "pg_check_arith": [ctypes.c_int, ctypes.c_uint64, _dp, _dp, _dp, _dp],
        "pg_check_bond": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _dp, _dp,
                          _dp, _dp],
        "pg_check_ordered_sum": [ctypes.c_int, ctypes.c_uint32, _u64p, ctypes.POINTER(ctypes.c_int32),
                                 _dp, _dp, _dp, _dp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


class PgError(RuntimeError):
    pass


def check(rc):
    if rc != PG_OK:
        msg = lib().pg_last_error().decode(errors="replace")
        raise PgError(f"libpg error {rc}: {msg}")


def device_count():
    n = ctypes.c_int()
    check(lib().pg_device_count(ctypes.byref(n)))
    return int(n.value)
