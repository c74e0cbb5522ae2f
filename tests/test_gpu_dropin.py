"""The drop-in executables (perigrain_amd/bin/simulate2d, simulate3d) on a
GPU: run them in a directory laid out like the reference's repository root
(config/main.conf, data/hdf5/all.h5) and check every output file against
the oracle stepped to the same snapshot steps.  tc/wall files must be
bit-identical to the oracle's state (2D without walls, 3D with walls)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from perigrain_amd import h5io, setups

pytestmark = pytest.mark.gpu


def _layout(tmp, s, p, **cfg):
    os.makedirs(tmp / "config", exist_ok=True)
    os.makedirs(tmp / "data" / "hdf5", exist_ok=True)
    h5io.write_all_h5(str(tmp / "data" / "hdf5" / "all.h5"), s)
    keys = h5io.wall_speed_keys(s.dim, p.wall_speed[: (4 if s.dim == 2 else 6)])
    keys.update(cfg)
    h5io.write_config(str(tmp / "config" / "main.conf"), p, **keys)


def _run(tmp, dim):
    exe = h5io.simulate_exe(dim)
    print(f"running {exe} in {tmp}", flush=True)
    r = subprocess.run([exe], cwd=str(tmp), stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:]
    return r.stdout


def _same(a, b, rtol, scale):
    if rtol == 0:
        return np.array_equal(a, b)
    return np.max(np.abs(a - b), initial=0.0) <= rtol * scale


def _check_against_oracle(tmp, s, p, counters, modulo, threads, start_setup=None, rtol=0.0):
    # rtol > 0: fields and reactions within rtol of their largest value (2D
    # walls, libm asin); Connectivity and wall positions always exact
    o = oracle.Oracle(start_setup or s, p, threads=threads)
    for c in counters:
        o.step(modulo)
        t = h5io.read_tc(str(tmp / "output" / "hdf5" / f"tc_{c:05d}.h5"), s.n_grains)
        ref = {k: o.read(k) for k in ("CurrPos", "vel", "acc", "force")}
        for g in range(s.n_grains):
            sl = s.grain_nodes(g)
            for k in ref:
                scale = max(np.max(np.abs(ref[k])), 1e-300)
                assert _same(t[g][k], ref[k][sl], rtol, scale), (c, g, k)
            assert np.array_equal(t[g]["Connectivity"], o.conn(g)), (c, g)
        wi, rr = h5io.read_wall(str(tmp / "output" / "hdf5" / f"wall_{c:05d}.h5"))
        owi, orr = o.wall()
        nside = 4 if s.dim == 2 else 6
        assert np.array_equal(wi, owi[:nside]), c
        assert _same(rr, orr[:nside], rtol, max(np.max(np.abs(orr)), 1e-300)), c
    o.close()


def test_simulate2d_two_disk_fracture(tmp_path, oracle_threads):
    s, p = setups.two_disk_collision(seed=5, gap=2e-5, speed=40.0)
    p.timesteps, p.modulo = 150, 50
    _layout(tmp_path, s, p)
    out = _run(tmp_path, 2)
    assert "Runtime:" in out
    info = h5io.read_plotinfo(str(tmp_path / "output" / "hdf5" / "plotinfo.h5"))
    assert list(info["f_l_counter"]) == [1, 3] and info["dimension"] == 2
    assert info["dt"] == p.dt and info["modulo"] == 50
    rt = h5io.read_dataset(str(tmp_path / "output" / "hdf5" / "run_time.h5"), "t_ind")[:, 0]
    assert list(rt) == [50.0, 100.0, 150.0]
    _check_against_oracle(tmp_path, s, p, [1, 2, 3], 50, oracle_threads)


def test_simulate3d_compression_walls_and_resume(tmp_path, oracle_threads):
    s, p = setups.compression_3d(seed=2, n_side=2, R=0.5e-3, h_div=5, cube=3e-3,
                                 wall_speed=-5.0)
    p.timesteps, p.modulo = 80, 40
    _layout(tmp_path, s, p)
    _run(tmp_path, 3)
    _check_against_oracle(tmp_path, s, p, [1, 2], 40, oracle_threads)

    # resume from counter 2 (timeloop.h:508-557): disp = CurrPos - pos, the
    # bonds from its Connectivity, the walls from wall_00002.h5
    h5io.write_config(str(tmp_path / "config" / "main.conf"), p,
                      **h5io.wall_speed_keys(3, p.wall_speed[:6]), do_resume=1, resume_ind=2,
                      wall_resume=1, timesteps=40)
    hd = tmp_path / "output" / "hdf5"
    cfg = h5io.read_config(str(tmp_path / "config" / "main.conf"))
    s2, p2 = h5io.load_setup(str(tmp_path / "data" / "hdf5" / "all.h5"), 3, cfg,
                             resume=(str(hd / "tc_00002.h5"), p.enable_fracture,
                                     str(hd / "wall_00002.h5")))
    _run(tmp_path, 3)
    info = h5io.read_plotinfo(str(hd / "plotinfo.h5"))
    assert list(info["f_l_counter"]) == [3, 3]
    _check_against_oracle(tmp_path, s, p2, [3], 40, oracle_threads, start_setup=s2)


@pytest.mark.parametrize("case", range(4))
def test_random_dropin_runs_match_the_oracle(case, tmp_path, oracle_threads):
    """Random small workloads with random switches (as the randomised parity
    tests draw them), written as all.h5 and config/main.conf and run through
    the executables with a random step count and snapshot interval: every
    tc/wall file against the oracle at its step."""
    from test_gpu_random import _draw_small
    rng = np.random.default_rng(9000 + case)
    dim = 3 if case % 2 else 2
    s, p, names = _draw_small(rng, dim)
    modulo = int(rng.integers(15, 45))
    n_snap = int(rng.integers(2, 5))
    p.timesteps, p.modulo = modulo * n_snap, modulo
    print(f"case {case}: {dim}D, {s.n_nodes} nodes, {p.timesteps} steps every {modulo}, "
          f"{'+'.join(names) or 'plain'}", flush=True)
    _layout(tmp_path, s, p)
    _run(tmp_path, dim)
    info = h5io.read_plotinfo(str(tmp_path / "output" / "hdf5" / "plotinfo.h5"))
    assert list(info["f_l_counter"]) == [1, n_snap]
    rtol = 1e-10 if (dim == 2 and s.allow_wall) else 0.0
    _check_against_oracle(tmp_path, s, p, list(range(1, n_snap + 1)), modulo, oracle_threads,
                          rtol=rtol)
