"""On-device post-processing (pg_grain_stats) against numpy on the fields
read back from the device: the per-particle quantities of the reference's
read_sim_output.py extract_bulk (:143-266) and read_force_val (:287-316)."""
import numpy as np
import pytest

from perigrain_amd import Stepper, setups

pytestmark = pytest.mark.gpu


def _numpy_stats(g, s):
    d = s.dim
    no = np.asarray(s.node_offset)
    F, v, a, x = (g.read(f) for f in ("force", "vel", "acc", "CurrPos"))
    # bonds per node at creation (conn2NArr: each row adds one to both ends)
    # and intact bonds now, from the per-slot status in NbdArr order
    deg = np.zeros(int(no[-1]), dtype=np.int64)
    co = np.asarray(s.conn_offset)
    conn = np.asarray(s.conn).reshape(-1, 2).astype(np.int64)
    for k in range(s.n_grains):
        c = conn[co[k]:co[k + 1]] + no[k]
        np.add.at(deg, c.ravel(), 1)
    intact = g.bond_intact().astype(np.int64)
    now = np.add.reduceat(intact, np.concatenate([[0], np.cumsum(deg)[:-1]]))
    now[deg == 0] = 0
    rows = []
    for k in range(s.n_grains):
        sl = slice(no[k], no[k + 1])
        vra = (s.vol[sl] * s.rho[k])[:, None] * a[sl]
        with np.errstate(invalid="ignore", divide="ignore"):
            dmg = np.mean((deg[sl] - now[sl]) / deg[sl])
        rows.append(np.concatenate([F[sl].sum(0), v[sl].mean(0), a[sl].mean(0), x[sl].mean(0),
                                    [dmg], vra.sum(0)]))
    return np.array(rows)


def _check(stats, ref, d):
    got = np.concatenate([stats["force_sum"], stats["mean_vel"], stats["mean_acc"],
                          stats["mean_CurrPos"], stats["damage"][:, None],
                          stats["vol_rho_acc_sum"]], axis=1)
    for c in range(got.shape[1]):
        scale = max(np.max(np.abs(ref[:, c])), 1e-300)
        assert np.max(np.abs(got[:, c] - ref[:, c])) <= 1e-12 * scale, f"column {c}"


def test_grain_stats_2d_fracture():
    s, p = setups.two_disk_collision(seed=4, gap=1e-5, speed=50.0)
    nodes = s.grain_nodes(0)
    s.vel[nodes] += 1e5 * (s.pos[nodes] - s.pos[nodes].mean(axis=0))
    with Stepper(s, p, device=0) as g:
        g.step(150)
        st = g.grain_stats()
        ref = _numpy_stats(g, s)
        assert np.all(st["damage"] >= 0) and st["damage"][0] > 0, "disk 0 should be damaged"
        _check(st, ref, 2)


def test_grain_stats_3d_compression():
    s, p = setups.compression_3d(seed=5, n_side=2, R=0.5e-3, h_div=5, cube=3e-3, wall_speed=-5.0)
    with Stepper(s, p, device=0) as g:
        g.step(120)
        _check(g.grain_stats(), _numpy_stats(g, s), 3)


def test_grain_stats_of_batched_instances():
    # each instance of a batched handle reports its own grains, as a
    # single-instance handle of the same setup does
    sp = [setups.settle_instance_2d(seed) for seed in (21, 22)]
    ss = [setups.subset(s, list(range(min(10, s.n_grains)))) for s, _ in sp]
    ps = [p for _, p in sp]
    with Stepper(ss, ps, device=0) as b:
        b.step(80)
        for k, (s, p) in enumerate(zip(ss, ps)):
            st = b.grain_stats(inst=k)
            assert st["force_sum"].shape == (s.n_grains, 2)
            with Stepper(s, p, device=0) as one:
                one.step(80)
                _check(st, _numpy_stats(one, s), 2)
