"""File side of the drop-in boundary: ctypes over libpgio (include/pgio.h).

Reads and writes the reference's files: config/main.conf
(src/read/read_config.h), data/hdf5/all.h5 (schema of exp_dict.py:705-770,
read by src/read/load_hdf5.h:41-201) and output/hdf5/{plotinfo,tc_%05u,
wall_%05u,run_time}.h5 (src/particle/timeloop.h:561-582,903-949,1010-1014).
h5py is not needed: libpgio uses the HDF5 C library.
"""
import ctypes
import os

import numpy as np

from . import _ffi
from .setups import Params, Setup

_HERE = os.path.dirname(os.path.abspath(__file__))
# PERIGRAIN_LIBPGIO selects another build of the library (e.g. one built with
# the address and undefined-behaviour sanitizers)
LIBPGIO_PATH = os.environ.get("PERIGRAIN_LIBPGIO") or os.path.join(_HERE, "libpgio.so")
BIN_DIR = os.path.join(_HERE, "bin")

F64, I64, U32, I32 = 0, 1, 2, 3
_NP = {F64: np.float64, I64: np.int64, U32: np.uint32, I32: np.int32}

_u32, _i32, _f64 = ctypes.c_uint32, ctypes.c_int32, ctypes.c_double


class PgioConfig(ctypes.Structure):
    """pgio_config: every config/main.conf key (read_config.h:14-120)."""
    _fields_ = [
        ("timesteps", _u32), ("modulo", _u32), ("dt", _f64), ("is_parallel", _i32),
        ("do_resume", _i32), ("wall_resume", _i32), ("resume_ind", _u32),
        ("save_file", _i32), ("enable_fracture", _i32), ("nl_bdry_only", _i32),
        ("allow_damping", _i32), ("allow_friction", _i32),
        ("normal_stiffness", _f64), ("damping_ratio", _f64), ("friction_coefficient", _f64),
        ("self_contact", _i32), ("self_contact_rad", _f64), ("enable_torque", _i32),
        ("set_movable_index", _i32), ("set_movable_timestep", _i32),
        ("set_stoppable_index", _i32), ("set_stoppable_timestep", _i32),
        ("reset_partzero_y", _i32), ("reset_partzero_y_timestep", _u32), ("wheel_rad", _f64),
        ("wall_left", _f64), ("wall_right", _f64), ("wall_top", _f64), ("wall_bottom", _f64),
        ("wall_x_min", _f64), ("wall_y_min", _f64), ("wall_z_min", _f64),
        ("wall_x_max", _f64), ("wall_y_max", _f64), ("wall_z_max", _f64),
        ("speed_wall_left", _f64), ("speed_wall_right", _f64), ("speed_wall_top", _f64),
        ("speed_wall_bottom", _f64),
        ("speed_wall_x_min", _f64), ("speed_wall_y_min", _f64), ("speed_wall_z_min", _f64),
        ("speed_wall_x_max", _f64), ("speed_wall_y_max", _f64), ("speed_wall_z_max", _f64),
        ("gradient_extforce", _i32), ("extforce_maxstep", _u32), ("unknown_keys", _u32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def lib():
    """Load libpgio.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPGIO_PATH):
        raise RuntimeError(f"libpgio.so not found at {LIBPGIO_PATH}; build it with "
                           "`python -m perigrain_amd.build`")
    L = ctypes.CDLL(LIBPGIO_PATH)
    L.pgio_last_error.restype = ctypes.c_char_p
    H = ctypes.c_void_p
    cp = ctypes.c_char_p
    vp = ctypes.c_void_p
    u64 = ctypes.c_uint64
    sig = {
        "pgio_config_defaults": [ctypes.POINTER(PgioConfig)],
        "pgio_read_config": [cp, ctypes.POINTER(PgioConfig)],
        "pgio_load_setup": [cp, ctypes.c_int, ctypes.POINTER(PgioConfig), ctypes.POINTER(H)],
        "pgio_setup_view": [H, ctypes.POINTER(_ffi.PgSetup), ctypes.POINTER(_ffi.PgParams)],
        "pgio_resume": [H, cp, ctypes.c_int, cp],
        "pgio_setup_free": [H],
        "pgio_write_plotinfo": [cp, ctypes.c_int, _u32, _u32, _f64, _u32, _i32, vp],
        "pgio_write_tc": [cp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp],
        "pgio_write_wall": [cp, ctypes.c_int, vp, vp],
        "pgio_write_runtime": [cp, u64, vp, vp],
        "pgio_h5_write": [cp, ctypes.c_int, cp, ctypes.c_int, ctypes.c_int, u64, u64, vp],
        "pgio_h5_shape": [cp, cp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(u64)],
        "pgio_h5_read": [cp, cp, ctypes.c_int, vp, u64],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


class PgioError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise PgioError(f"libpgio error {rc}: {lib().pgio_last_error().decode(errors='replace')}")


def _b(s):
    return os.fsencode(s)


# ------------------------------------------------------------- generic access

def write_dataset(path, name, data, dtype=None, rank=None, create_new=False):
    """Write one dataset (missing groups are created)."""
    a = np.asarray(data)
    if dtype is not None:
        a = a.astype(dtype)
    t = {np.dtype(np.float64): F64, np.dtype(np.int64): I64, np.dtype(np.uint32): U32,
         np.dtype(np.int32): I32}[a.dtype]
    a = np.ascontiguousarray(a)
    r = rank or max(1, a.ndim)
    rows = a.shape[0] if a.ndim >= 1 else 1
    cols = a.shape[1] if a.ndim >= 2 else 1
    _check(lib().pgio_h5_write(_b(path), int(create_new), _b(name), t, r, rows, cols,
                               a.ctypes.data_as(ctypes.c_void_p) if a.size else None))


def dataset_shape(path, name):
    rank = ctypes.c_int()
    dims = (ctypes.c_uint64 * 2)()
    _check(lib().pgio_h5_shape(_b(path), _b(name), ctypes.byref(rank), dims))
    return tuple(int(dims[k]) for k in range(rank.value))


def read_dataset(path, name, t=F64):
    shape = dataset_shape(path, name)
    out = np.empty(shape, dtype=_NP[t])
    _check(lib().pgio_h5_read(_b(path), _b(name), t,
                              out.ctypes.data_as(ctypes.c_void_p) if out.size else None,
                              out.size))
    return out


# ------------------------------------------------------------------- config

def default_config():
    c = PgioConfig()
    _check(lib().pgio_config_defaults(ctypes.byref(c)))
    return c


def read_config(path):
    c = default_config()
    _check(lib().pgio_read_config(_b(path), ctypes.byref(c)))
    return c


def write_config(path, params: Params, **extra):
    """Write config/main.conf for `params` (plus any extra keys)."""
    keys = {
        "timesteps": params.timesteps, "modulo": params.modulo, "dt": repr(float(params.dt)),
        "enable_fracture": params.enable_fracture, "nl_bdry_only": params.nl_bdry_only,
        "allow_damping": params.allow_damping, "allow_friction": params.allow_friction,
        "self_contact": params.self_contact, "gradient_extforce": params.gradient_extforce,
        "extforce_maxstep": params.extforce_maxstep, "enable_torque": params.enable_torque,
        "set_movable_index": params.set_movable_index,
        "set_movable_timestep": params.set_movable_timestep,
        "set_stoppable_index": params.set_stoppable_index,
        "set_stoppable_timestep": params.set_stoppable_timestep,
    }
    keys.update(extra)
    with open(path, "w") as f:
        f.write("# written by perigrain_amd.h5io.write_config\n")
        for k, v in keys.items():
            f.write(f"{k} = {v}\n")


def wall_speed_keys(dim, speeds):
    names = (("speed_wall_left", "speed_wall_right", "speed_wall_top", "speed_wall_bottom")
             if dim == 2 else
             ("speed_wall_x_min", "speed_wall_y_min", "speed_wall_z_min",
              "speed_wall_x_max", "speed_wall_y_max", "speed_wall_z_max"))
    return {n: repr(float(v)) for n, v in zip(names, speeds)}


# ------------------------------------------------------------------ all.h5

def write_all_h5(path, s: Setup, int_dtype=np.int64):
    """Write `s` in the reference's all.h5 layout (exp_dict.py:705-770):
    1-based P_%05d groups, [[scalar]] datasets, int64 index arrays."""
    d = s.dim
    write_dataset(path, "total_particles_univ", [[s.n_grains]], int_dtype, create_new=True)
    for g in range(s.n_grains):
        grp = f"P_{g + 1:05d}/"
        sl = s.grain_nodes(g)
        c0, c1 = int(s.conn_offset[g]), int(s.conn_offset[g + 1])
        b0, b1 = int(s.bdry_offset[g]), int(s.bdry_offset[g + 1])
        k0, k1 = int(s.clamped_offset[g]), int(s.clamped_offset[g + 1])
        write_dataset(path, grp + "Pos", s.pos[sl].reshape(-1, d), np.float64)
        write_dataset(path, grp + "Vol", s.vol[sl].reshape(-1, 1), np.float64)
        write_dataset(path, grp + "Connectivity", np.asarray(s.conn[c0:c1]).reshape(-1, 2),
                      int_dtype)
        write_dataset(path, grp + "bdry_nodes", np.asarray(s.bdry_nodes[b0:b1]).reshape(-1, 1),
                      int_dtype)
        # rank-1 float64, possibly empty, as the generator stores it
        write_dataset(path, grp + "clamped_nodes",
                      np.asarray(s.clamped_nodes[k0:k1], dtype=np.float64), rank=1)
        for name in ("disp", "vel", "acc", "extforce"):
            write_dataset(path, grp + name, getattr(s, name)[sl].reshape(-1, d), np.float64)
        for name in ("delta", "rho", "cnot", "snot"):
            write_dataset(path, grp + name, [[float(getattr(s, name)[g])]], np.float64)
        for name in ("movable", "breakable", "stoppable"):
            write_dataset(path, grp + name, [[int(getattr(s, name)[g])]], int_dtype)
        write_dataset(path, grp + "torque_axis", [[int(s.torque_axis[g])]], int_dtype)
        write_dataset(path, grp + "torque_val", [[float(s.torque_val[g])]], np.float64)
    for name in ("contact_radius", "normal_stiffness", "damping_ratio", "friction_coefficient"):
        write_dataset(path, "pairwise/" + name, [[float(getattr(s, name))]], np.float64)
    write_dataset(path, "wall/allow_wall", [[int(s.allow_wall)]], int_dtype)
    nside = 4 if d == 2 else 6
    write_dataset(path, "wall/geom_wall_info", np.asarray(s.wall[:nside]).reshape(-1, 1),
                  np.float64)


def _copy(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def load_setup(all_h5, dim, config=None, resume=None):
    """Load all.h5 with `config` (a PgioConfig) applied, as libpgio does for
    the drop-in executables; returns (Setup, Params).  resume =
    (tc_h5, load_connectivity, wall_h5 or None) applies a snapshot."""
    cfg = config if config is not None else default_config()
    h = ctypes.c_void_p()
    _check(lib().pgio_load_setup(_b(all_h5), dim, ctypes.byref(cfg), ctypes.byref(h)))
    try:
        if resume is not None:
            tc, lc, wall = resume
            _check(lib().pgio_resume(h, _b(tc), int(lc), _b(wall) if wall else None))
        st = _ffi.PgSetup()
        pp = _ffi.PgParams()
        _check(lib().pgio_setup_view(h, ctypes.byref(st), ctypes.byref(pp)))
        G = st.n_grains
        no = _copy(st.node_offset, G + 1, np.int64)
        N = int(no[-1])
        co = _copy(st.conn_offset, G + 1, np.int64)
        bo = _copy(st.bdry_offset, G + 1, np.int64)
        ko = _copy(st.clamped_offset, G + 1, np.int64)
        s = Setup(
            dim=st.dim, node_offset=no,
            pos=_copy(st.pos, N * dim, np.float64).reshape(N, dim),
            vol=_copy(st.vol, N, np.float64),
            disp=_copy(st.disp, N * dim, np.float64).reshape(N, dim),
            vel=_copy(st.vel, N * dim, np.float64).reshape(N, dim),
            acc=_copy(st.acc, N * dim, np.float64).reshape(N, dim),
            extforce=_copy(st.extforce, N * dim, np.float64).reshape(N, dim),
            conn_offset=co, conn=_copy(st.conn, 2 * int(co[-1]), np.uint32).reshape(-1, 2),
            bdry_offset=bo, bdry_nodes=_copy(st.bdry_nodes, int(bo[-1]), np.uint32),
            clamped_offset=ko, clamped_nodes=_copy(st.clamped_nodes, int(ko[-1]), np.uint32),
            delta=_copy(st.delta, G, np.float64), rho=_copy(st.rho, G, np.float64),
            cnot=_copy(st.cnot, G, np.float64), snot=_copy(st.snot, G, np.float64),
            movable=_copy(st.movable, G, np.int32), breakable=_copy(st.breakable, G, np.int32),
            stoppable=_copy(st.stoppable, G, np.int32),
            torque_axis=_copy(st.torque_axis, G, np.uint32),
            torque_val=_copy(st.torque_val, G, np.float64),
            contact_radius=st.contact_radius, normal_stiffness=st.normal_stiffness,
            damping_ratio=st.damping_ratio, friction_coefficient=st.friction_coefficient,
            allow_wall=st.allow_wall, wall=np.array([st.wall[k] for k in range(6)]))
        p = Params(
            dt=pp.dt, timesteps=cfg.timesteps, modulo=cfg.modulo,
            enable_fracture=pp.enable_fracture, nl_bdry_only=pp.nl_bdry_only,
            allow_damping=pp.allow_damping, allow_friction=pp.allow_friction,
            self_contact=pp.self_contact, gradient_extforce=pp.gradient_extforce,
            extforce_maxstep=pp.extforce_maxstep, enable_torque=pp.enable_torque,
            set_movable_index=pp.set_movable_index, set_movable_timestep=pp.set_movable_timestep,
            set_stoppable_index=pp.set_stoppable_index,
            set_stoppable_timestep=pp.set_stoppable_timestep,
            reset_partzero_y=pp.reset_partzero_y,
            reset_partzero_y_timestep=pp.reset_partzero_y_timestep, wheel_rad=pp.wheel_rad,
            wall_speed=np.array([pp.wall_speed[k] for k in range(6)]))
        return s, p
    finally:
        lib().pgio_setup_free(h)


# ------------------------------------------------------------------ outputs

def read_tc(path, n_grains):
    """Per grain: dict with CurrPos, vel, acc, force (float64) and
    Connectivity (uint32)."""
    out = []
    for g in range(n_grains):
        grp = f"P_{g:05d}/"
        d = {k: read_dataset(path, grp + k) for k in ("CurrPos", "vel", "acc", "force")}
        d["Connectivity"] = read_dataset(path, grp + "Connectivity", U32)
        out.append(d)
    return out


def read_wall(path):
    return read_dataset(path, "wall_info")[:, 0], read_dataset(path, "reaction")


def read_plotinfo(path):
    return {
        "f_l_counter": read_dataset(path, "f_l_counter", U32)[:, 0],
        "dimension": int(read_dataset(path, "dimension", U32)[0, 0]),
        "dt": float(read_dataset(path, "dt")[0, 0]),
        "modulo": int(read_dataset(path, "modulo", U32)[0, 0]),
        "allow_wall": int(read_dataset(path, "/wall/allow_wall", I32)[0, 0]),
        "geom_wall_info": read_dataset(path, "/wall/geom_wall_info")[:, 0],
    }


def simulate_exe(dim):
    exe = os.path.join(BIN_DIR, f"simulate{dim}d")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} not built; run `python -m perigrain_amd.build`")
    return exe
