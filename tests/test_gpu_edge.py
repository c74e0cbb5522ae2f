"""GPU parity on degenerate inputs, each against the oracle bit for bit:
grains of one node, grains whose nodes have no bonds, a chain grain of 65
nodes (rows of two to six bonds across a 64-row slice boundary), a handle
with no bonds at all, a 3D handle with a one-node grain between spheres and
walls, whole runs of the range-tested bond loop (a reference coordinate or a
displacement below 2^-240), and the 2D cases as instances of one batched
handle.  The reference runs such grains like any other (timeloop.h:608-878
loops over each grain's nodes and NbdArr rows, however short)."""
import numpy as np
import pytest

import oracle
from perigrain_amd import Stepper, setups
from test_gpu_parity import FIELDS, _advance, _assert_close, _assert_exact, _both

pytestmark = pytest.mark.gpu


def _points_grain(pts, h, dim, delta, mat, bonds=True):
    """Grain of the given nodes; bonds between nodes within delta unless
    bonds is False; every node is a boundary node."""
    pts = np.asarray(pts, dtype=float).reshape(-1, dim)
    n = len(pts)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n)
             if bonds and np.linalg.norm(pts[j] - pts[i]) <= delta]
    conn = np.array(pairs, dtype=np.uint32).reshape(-1, 2)
    rho, c, s0 = mat(delta)
    return setups.Grain(pts, np.full(n, h ** dim), conn, np.arange(n, dtype=np.uint32), delta,
                        rho, c, s0, float(np.max(np.linalg.norm(pts - pts.mean(axis=0), axis=1))))


def _mixed_2d(seed):
    rng = np.random.default_rng(seed)
    R = 1e-3
    h = R / 12
    delta = 3 * h
    Rc = delta / 5
    disk = setups.make_grain_2d("disk", R, h, rng, Rc)
    one = _points_grain([[0.0, 0.0]], h, 2, delta, setups.material_2d)
    loose = _points_grain([[0.0, 0.0], [4 * h, 0.0], [0.0, 4 * h]], h, 2, delta,
                          setups.material_2d)  # nodes beyond delta: no bonds
    chain = _points_grain(np.stack([np.arange(65) * h, rng.uniform(-0.05, 0.05, 65) * h], 1),
                          h, 2, delta, setups.material_2d)
    grains = [disk, one, loose, chain]
    lo, hi = disk.pos.min(axis=0), disk.pos.max(axis=0)
    gap = 1.5 * Rc  # every grain starts just outside contact range of the disk
    centers = [np.zeros(2), np.array([hi[0] + gap, 0.0]),
               np.array([lo[0] - gap - 4 * h, 0.3 * R]), np.array([-32 * h, hi[1] + gap])]
    vels = [np.array([5.0, 0.0]), np.array([-40.0, 3.0]), np.array([35.0, -5.0]),
            np.array([0.0, -30.0])]
    s = setups.assemble(2, grains, centers, vels, Rc, disk.cnot, 0.8, 0.3, 0.0, 0,
                        [-4 * R, 4 * R, 2 * R, -2 * R])
    p = setups.Params(dt=1e-8, timesteps=400, modulo=100, enable_fracture=1)
    return s, p


def _no_bonds_2d(seed):
    rng = np.random.default_rng(seed)
    h = 1e-4
    delta = 3 * h
    Rc = delta / 5
    grains = [_points_grain([[0.0, 0.0]], h, 2, delta, setups.material_2d) for _ in range(6)]
    centers = [np.array([k * (Rc + 2e-6), 0.0]) for k in range(6)]
    vels = [rng.normal(size=2) * 20.0 for _ in range(6)]
    s = setups.assemble(2, grains, centers, vels, Rc, grains[0].cnot, 0.8, 0.3, 0.0, 0,
                        [-1e-2, 1e-2, 1e-2, -1e-2])
    p = setups.Params(dt=1e-8, timesteps=400, modulo=100, enable_fracture=1)
    return s, p


def _append_node_grain(s, x, v, h, mat):
    """Append a grain of one node at x with velocity v (a copy of the last
    grain's gravity), volume h^dim and the material mat(3h)."""
    g = _points_grain(np.zeros((1, s.dim)), h, s.dim, 3 * h, mat)
    s.node_offset = np.append(s.node_offset, s.node_offset[-1] + 1)
    s.conn_offset = np.append(s.conn_offset, s.conn_offset[-1])
    s.bdry_offset = np.append(s.bdry_offset, s.bdry_offset[-1] + 1)
    s.bdry_nodes = np.append(s.bdry_nodes, np.uint32(0))
    s.clamped_offset = np.append(s.clamped_offset, s.clamped_offset[-1])
    s.pos = np.concatenate([s.pos, np.asarray(x, float).reshape(1, -1)])
    s.vel = np.concatenate([s.vel, np.asarray(v, float).reshape(1, -1)])
    s.vol = np.append(s.vol, h ** s.dim)
    s.disp = np.zeros_like(s.pos)
    s.acc = np.zeros_like(s.pos)
    s.extforce = np.concatenate([s.extforce, s.extforce[-1:]])
    for f, val in (("delta", g.delta), ("rho", g.rho), ("cnot", g.cnot), ("snot", g.snot)):
        setattr(s, f, np.append(getattr(s, f), val))
    for f in ("movable", "breakable", "stoppable"):
        setattr(s, f, np.append(getattr(s, f), np.int32(1)))
    s.torque_axis = np.append(s.torque_axis, np.uint32(2))
    s.torque_val = np.append(s.torque_val, 0.0)


def _one_node_3d(seed):
    s, p = setups.compression_3d(seed=seed, n_side=2, R=0.5e-3, h_div=4, cube=3e-3,
                                 wall_speed=-10.0)
    lo = s.pos.min(axis=0)
    # the node starts 1.5 contact radii above the bottom wall, under the
    # lowest sphere, and falls onto the wall before rising into the spheres
    z0 = s.wall[2] + 1.5 * s.contact_radius
    G = s.n_grains
    _append_node_grain(s, [lo[0] + 0.3e-3, lo[1] + 0.3e-3, z0], [10.0, 10.0, -15.0],
                       0.5e-3 / 4, setups.material_3d)
    assert s.n_grains == G + 1
    return s, p


def _free_point(s):
    """The point inside the walls (4 contact radii in) farthest from every node."""
    from scipy.spatial import cKDTree
    d = s.dim
    if d == 3:
        lo, hi = np.array(s.wall[:3]), np.array(s.wall[3:6])
    else:  # left, right, top, bottom
        lo, hi = np.array([s.wall[0], s.wall[3]]), np.array([s.wall[1], s.wall[2]])
    lo, hi = lo + 4 * s.contact_radius, hi - 4 * s.contact_radius
    grid = np.stack(np.meshgrid(*[np.linspace(lo[k], hi[k], 25 if d == 3 else 200)
                                  for k in range(d)], indexing="ij"), -1).reshape(-1, d)
    dist, _ = cKDTree(s.pos).query(grid)
    assert np.max(dist) > 4 * s.contact_radius
    return grid[int(np.argmax(dist))]


def _nice(x):
    m = np.abs(x)
    return (m == 0.0) | ((m >= 2.0 ** -240) & (m <= 2.0 ** 240))


def _checked_2d(seed):
    """The two-disk fracture with one node's y moved to 1e-80 (from within
    0.05h of zero): a reference coordinate that is not nice, so every step
    takes the bond loop with the per-bond range test and IEEE fallback."""
    s, p = setups.two_disk_collision(seed=seed, gap=1e-5, speed=50.0)
    nodes = s.grain_nodes(0)
    s.vel[nodes] += 1e5 * (s.pos[nodes] - s.pos[nodes].mean(axis=0))
    k = int(np.argmin(np.abs(s.pos[:, 1])))
    s.pos[k, 1] = 1e-80
    assert not _nice(s.pos).all()
    return s, p


def _checked_3d(seed):
    """The 8-sphere compression with one more one-node grain in the widest
    free space, moving at 1e-75 m/s along x with no external force: its
    displacement stays below 2^-240, so every step after the first takes the
    range-tested bond loop."""
    s, p = setups.compression_3d(seed=seed, n_side=2, R=0.5e-3, h_div=4, cube=3e-3,
                                 wall_speed=-10.0)
    _append_node_grain(s, _free_point(s), [1e-75, 0.0, 0.0], 0.5e-3 / 4, setups.material_3d)
    s.extforce[-1] = 0.0
    return s, p


def _checked_2d_large(seed):
    """The same free node in a 250-grain 2D settle (76k nodes, one lane per
    node) with grains thrown at 60 m/s: the range-tested loop of the
    large-workload kernel over a whole run."""
    s, p = setups.bulk_settle_2d(seed=seed, n_grains=250, box=(40e-3, 80e-3),
                                 nodes_per_grain=300)
    rng = np.random.default_rng(seed)
    for g in range(s.n_grains):
        s.vel[s.grain_nodes(g)] += rng.normal(size=2) * 60.0
    _append_node_grain(s, _free_point(s), [1e-75, 0.0], 1e-4, setups.material_2d)
    s.extforce[-1] = 0.0
    return s, p


CASES = {"2d-one-node-loose-and-chain-grains": _mixed_2d, "2d-no-bonds": _no_bonds_2d,
         "3d-one-node-grain": _one_node_3d, "2d-range-tested-loop": _checked_2d,
         "3d-range-tested-loop": _checked_3d, "2d-range-tested-loop-76k": _checked_2d_large}


@pytest.mark.parametrize("name", sorted(CASES))
def test_degenerate_grains_bit_exact(name, oracle_threads):
    s, p = CASES[name](7)
    print(f"{name}: {s.n_nodes} nodes, {s.n_grains} grains, {len(s.conn)} bonds", flush=True)
    g, o = _both(s, p, oracle_threads)
    try:
        done = 0
        for t in (1, 50, 200, 400):
            _advance(g, o, t - done)
            done = t
            if s.dim == 2 and s.allow_wall:  # libm asin in the 2D wall force
                _assert_close(g, o, 1e-10, f"{name} t={t}")
            else:
                _assert_exact(g, o, f"{name} t={t}")
        print(f"{name}: {len(o.contacts())} contact pairs, "
              f"{int((o.bond_intact() == 0).sum())} broken slots at t=400", flush=True)
        assert len(o.contacts()) > 0 or name == "2d-no-bonds"
        if name in ("3d-range-tested-loop", "2d-range-tested-loop-76k"):
            # the free node's x displacement is not nice
            ux = o.read("disp")[-1, 0]
            assert 0.0 < ux < 2.0 ** -240 and g.read("disp")[-1, 0] == ux
    finally:
        g.close()
        o.close()


def test_degenerate_instances_in_one_batched_handle(oracle_threads):
    """The two degenerate 2D workloads and a two-disk fracture in one handle
    (instances of 6, 506 and 1,394 nodes, one of them with no bonds): each
    instance against the oracle's run of it, bit for bit."""
    frac, pf = setups.two_disk_collision(seed=4, gap=1e-5, speed=50.0)
    nodes = frac.grain_nodes(0)
    frac.vel[nodes] += 1e5 * (frac.pos[nodes] - frac.pos[nodes].mean(axis=0))
    sp = [_no_bonds_2d(7), _mixed_2d(8), (frac, pf)]
    print(f"instances: {[s.n_nodes for s, _ in sp]} nodes", flush=True)
    os_ = [oracle.Oracle(s, p, threads=oracle_threads) for s, p in sp]
    try:
        with Stepper([s for s, _ in sp], [p for _, p in sp], device=0) as b:
            done = 0
            for t in (1, 100, 300):
                b.step(t - done)
                for k, o in enumerate(os_):
                    o.step(t - done, record_contacts=True)
                    for f in FIELDS:
                        a, c = b.read(f, inst=k), o.read(f)
                        assert np.array_equal(a, c), \
                            f"instance {k} t={t} {f}: {int((a != c).sum())} entries differ"
                    assert np.array_equal(b.bond_intact(inst=k), o.bond_intact()), \
                        f"instance {k} t={t} bond status"
                    assert np.array_equal(b.contacts(inst=k), o.contacts()), \
                        f"instance {k} t={t} contact pairs"
                done = t
            assert int((os_[2].bond_intact() == 0).sum()) > 100, "instance 2 should break bonds"
    finally:
        for o in os_:
            o.close()
