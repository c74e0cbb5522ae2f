"""File side of the drop-in boundary (libpgio): config/main.conf parsing,
all.h5 loading with the config overrides, resume, and the output writers.
Runs on the CPU (no GPU calls)."""
import os

import numpy as np
import pytest

from perigrain_amd import h5io, setups


def test_config_defaults_match_reference():
    # ConfigVal::ConfigVal(), src/read/read_config.h:61-120
    c = h5io.default_config().as_dict()
    assert (c["timesteps"], c["modulo"], c["dt"]) == (8000, 100, 0.02 / 1e5)
    assert (c["do_resume"], c["wall_resume"], c["resume_ind"], c["save_file"]) == (0, 0, 50, 1)
    assert (c["enable_fracture"], c["nl_bdry_only"], c["allow_damping"], c["allow_friction"]) == (
        0, 0, 1, 1)
    assert (c["normal_stiffness"], c["damping_ratio"], c["friction_coefficient"]) == (-1, -1, -1)
    assert c["self_contact"] == 1 and c["self_contact_rad"] == -1
    assert (c["set_movable_index"], c["set_movable_timestep"]) == (-1, -1)
    for k in ("wall_left", "wall_right", "wall_top", "wall_bottom", "wall_x_min", "wall_z_max"):
        assert c[k] == -999
    for k in ("speed_wall_left", "speed_wall_z_max"):
        assert c[k] == 0


def test_config_parse_whitespace_comments_and_unknown(tmp_path, capfd):
    p = tmp_path / "main.conf"
    p.write_text("# comment\n\n timesteps = 1 2 0\nmodulo=40\n  dt = 2e-8 \n"
                 "enable_fracture = 1\n#dt = 5\nnormal_stiffness = 3.5e17\n"
                 "wall_top = 0.25\nspeed_wall_z_max = -0.5\nnot_a_key = 3\n"
                 "set_movable_index = 2\nset_movable_timestep=7\n")
    c = h5io.read_config(str(p))
    err = capfd.readouterr().err
    assert "Wrong config name: not_a_key" in err
    assert c.unknown_keys == 1
    assert c.timesteps == 120  # whitespace inside the value is removed too
    assert c.modulo == 40 and c.dt == 2e-8 and c.enable_fracture == 1
    assert c.normal_stiffness == 3.5e17 and c.wall_top == 0.25 and c.speed_wall_z_max == -0.5
    assert (c.set_movable_index, c.set_movable_timestep) == (2, 7)


def test_missing_config_keeps_defaults(tmp_path, capfd):
    c = h5io.read_config(str(tmp_path / "nope.conf"))
    assert "Couldn't open config file" in capfd.readouterr().err
    assert c.timesteps == 8000


def _same_setup(a, b):
    for name in ("node_offset", "pos", "vol", "disp", "vel", "acc", "extforce", "conn_offset",
                 "conn", "bdry_offset", "bdry_nodes", "clamped_offset", "clamped_nodes",
                 "delta", "rho", "cnot", "snot", "movable", "breakable", "stoppable"):
        x, y = np.asarray(getattr(a, name)), np.asarray(getattr(b, name))
        assert x.shape == y.shape, name
        assert np.array_equal(x.astype(y.dtype), y), name
    for name in ("contact_radius", "normal_stiffness", "damping_ratio", "friction_coefficient",
                 "allow_wall"):
        assert getattr(a, name) == getattr(b, name), name


@pytest.mark.parametrize("int_dtype", [np.int64, np.int32, np.uint32])
@pytest.mark.parametrize("dim", [2, 3])
def test_all_h5_round_trip(tmp_path, dim, int_dtype):
    # int64 is what h5py writes from numpy's default integers (exp_dict.py);
    # int32 and uint32 files load the same through HDF5's conversions
    if dim == 2:
        s, _ = setups.two_disk_collision(seed=3)
    else:
        s, _ = setups.compression_3d(seed=3, n_side=2, R=0.5e-3, h_div=4, cube=3e-3)
    # a clamped node list and a non-trivial boundary list exercise the
    # float64 / int64 -> uint32 conversions of load_rank1 and load_col
    s.clamped_nodes = np.array([0, 5, 7], dtype=np.uint32)
    s.clamped_offset = np.zeros(s.n_grains + 1, dtype=np.int64)
    s.clamped_offset[1:] = 3
    s.clamped_offset[0] = 0
    path = str(tmp_path / "all.h5")
    h5io.write_all_h5(path, s, int_dtype=int_dtype)
    assert h5io.dataset_shape(path, "P_00001/Pos") == (int(s.node_offset[1]), dim)
    assert h5io.dataset_shape(path, "P_00001/clamped_nodes") == (3,)
    back, p = h5io.load_setup(path, dim)
    _same_setup(s, back)
    nside = 4 if dim == 2 else 6
    assert np.array_equal(back.wall[:nside], s.wall[:nside])
    # defaults of the config: nothing overridden, no wall motion
    assert p.dt == 0.02 / 1e5 and not p.enable_fracture
    assert np.all(p.wall_speed == 0)


def test_config_overrides_2d(tmp_path):
    s, _ = setups.two_disk_collision(seed=0)
    s.allow_wall = 1
    s.wall = np.array([-1.0, 1.0, 2.0, -2.0, 0, 0])
    path = str(tmp_path / "all.h5")
    h5io.write_all_h5(path, s)
    c = h5io.default_config()
    c.normal_stiffness = 7.0
    c.damping_ratio = 0.25
    c.wall_left = -3.0
    c.speed_wall_top = -0.5
    c.dt = 1e-8
    c.enable_fracture = 1
    back, p = h5io.load_setup(path, 2, c)
    assert back.normal_stiffness == 7.0 and back.damping_ratio == 0.25
    assert back.friction_coefficient == s.friction_coefficient  # -1 keeps the file value
    assert list(back.wall[:4]) == [-3.0, 1.0, 2.0, -2.0]  # left, right, top, bottom
    assert list(p.wall_speed[:4]) == [0.0, 0.0, -0.5, 0.0]
    assert p.dt == 1e-8 and p.enable_fracture == 1


def test_config_overrides_3d_quirk(tmp_path):
    # RectWall<3>::apply_config (contact.h:254-270): the 2D keys wall_top and
    # wall_right set z_max and y_max, before the 3D keys are applied
    s, _ = setups.compression_3d(seed=0, n_side=2, R=0.5e-3, h_div=4, cube=3e-3)
    path = str(tmp_path / "all.h5")
    h5io.write_all_h5(path, s)
    c = h5io.default_config()
    c.wall_top = 9.0
    c.wall_right = 8.0
    c.wall_y_max = 7.0
    c.speed_wall_z_max = -2.0
    back, p = h5io.load_setup(path, 3, c)
    w = back.wall
    assert w[5] == 9.0  # z_max from wall_top
    assert w[4] == 7.0  # y_max: wall_right, then wall_y_max
    assert list(w[:4]) == list(s.wall[:4])
    assert p.wall_speed[5] == -2.0


def test_resume_sets_disp_from_currpos(tmp_path):
    s, _ = setups.two_disk_collision(seed=0)
    path = str(tmp_path / "all.h5")
    h5io.write_all_h5(path, s)
    rng = np.random.default_rng(0)
    N, d = s.n_nodes, s.dim
    cur = s.pos + rng.normal(scale=1e-6, size=(N, d))
    vel = rng.normal(size=(N, d))
    acc = rng.normal(size=(N, d))
    frc = rng.normal(size=(N, d))
    # Connectivity as NbdArr2conn would write it: keep every other row
    conn_off = [0]
    rows = []
    for g in range(s.n_grains):
        c = s.conn[s.conn_offset[g]:s.conn_offset[g + 1]][::2]
        rows.append(c)
        conn_off.append(conn_off[-1] + len(c))
    conn = np.concatenate(rows).astype(np.uint32)
    tc = str(tmp_path / "tc_00004.h5")
    L = h5io.lib()
    import ctypes

    def p(a):
        return np.ascontiguousarray(a).ctypes.data_as(ctypes.c_void_p)

    no = np.ascontiguousarray(s.node_offset, dtype=np.int64)
    co = np.asarray(conn_off, dtype=np.int64)
    h5io._check(L.pgio_write_tc(os.fsencode(tc), d, s.n_grains, p(no), p(cur), p(vel), p(acc),
                                p(frc), p(co), p(conn)))
    wall_info = np.array([-1.0, 1.0, 2.0, -2.0])
    react = rng.normal(size=(4, d))
    wf = str(tmp_path / "wall_00004.h5")
    h5io._check(L.pgio_write_wall(os.fsencode(wf), d, p(wall_info), p(react)))
    back, _ = h5io.load_setup(path, 2, resume=(tc, 1, wf))
    assert np.array_equal(back.disp, cur - s.pos)  # disp = CurrPos - pos, bit for bit
    assert np.array_equal(back.vel, vel) and np.array_equal(back.acc, acc)
    assert np.array_equal(back.conn, conn.reshape(-1, 2))
    assert np.array_equal(back.conn_offset, co)
    assert list(back.wall[:4]) == list(wall_info)
    # without Connectivity the bonds of all.h5 stay
    back2, _ = h5io.load_setup(path, 2, resume=(tc, 0, None))
    assert np.array_equal(back2.conn, s.conn)
    # the tc file reads back per grain
    t = h5io.read_tc(tc, s.n_grains)
    sl = s.grain_nodes(1)
    assert np.array_equal(t[1]["CurrPos"], cur[sl]) and np.array_equal(t[1]["force"], frc[sl])
    assert np.array_equal(t[0]["Connectivity"], rows[0])
    wi, rr = h5io.read_wall(wf)
    assert np.array_equal(wi, wall_info) and np.array_equal(rr, react)


def test_plotinfo_and_runtime_writers(tmp_path):
    import ctypes
    L = h5io.lib()
    wall = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 5.0])
    pi = str(tmp_path / "plotinfo.h5")
    h5io._check(L.pgio_write_plotinfo(os.fsencode(pi), 3, 4, 9, 2e-8, 25, 1,
                                      wall.ctypes.data_as(ctypes.c_void_p)))
    info = h5io.read_plotinfo(pi)
    assert list(info["f_l_counter"]) == [4, 9] and info["dimension"] == 3
    assert info["dt"] == 2e-8 and info["modulo"] == 25 and info["allow_wall"] == 1
    assert np.array_equal(info["geom_wall_info"], wall)
    assert h5io.dataset_shape(pi, "f_l_counter") == (2, 1)
    rt = str(tmp_path / "run_time.h5")
    a = np.array([10.0, 12.0])
    b = np.array([100.0, 200.0])
    h5io._check(L.pgio_write_runtime(os.fsencode(rt), 2, a.ctypes.data_as(ctypes.c_void_p),
                                     b.ctypes.data_as(ctypes.c_void_p)))
    assert np.array_equal(h5io.read_dataset(rt, "run_time")[:, 0], a)
    assert np.array_equal(h5io.read_dataset(rt, "t_ind")[:, 0], b)


def test_missing_dataset_is_an_error(tmp_path):
    path = str(tmp_path / "all.h5")
    h5io.write_dataset(path, "total_particles_univ", [[1]], np.int64, create_new=True)
    with pytest.raises(h5io.PgioError, match="P_00001/Pos"):
        h5io.load_setup(path, 2)
