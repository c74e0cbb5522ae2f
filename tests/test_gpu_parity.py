"""GPU parity: the HIP stepper (through the libpg C ABI) against the oracle.

Bit-exact: bond status, broken-bond sets, contact pair lists, and every
FP64 field where no libm transcendental is involved (2D walls call asin,
whose last bit may differ between the host libm and the device's).
"""
import numpy as np
import pytest

import oracle
from perigrain_amd import Stepper, _ffi, setups

pytestmark = pytest.mark.gpu

FIELDS = ("CurrPos", "vel", "acc", "force")


def _both(s, p, threads):
    g = Stepper(s, p, device=0)
    o = oracle.Oracle(s, p, threads=threads)
    return g, o


def _advance(g, o, n):
    g.step(n)
    o.step(n, record_contacts=True)


def _assert_exact(g, o, what=""):
    for f in FIELDS:
        a, b = g.read(f), o.read(f)
        bad = np.argwhere(a != b)
        assert len(bad) == 0, f"{what} {f}: {len(bad)} entries differ, first {bad[:3].tolist()} " \
                              f"gpu={a[tuple(bad[0])]!r} oracle={b[tuple(bad[0])]!r}"
    assert np.array_equal(g.bond_intact(), o.bond_intact()), f"{what} bond status differs"
    assert np.array_equal(g.contacts(), o.contacts()), f"{what} contact pairs differ"
    wg, rg = g.wall()
    wo, ro = o.wall()
    assert np.array_equal(wg, wo), f"{what} wall positions differ"
    assert np.array_equal(rg, ro), f"{what} wall reaction differs: {rg} vs {ro}"


def _assert_close(g, o, rtol, what=""):
    for f in FIELDS:
        a, b = g.read(f), o.read(f)
        scale = max(np.max(np.abs(b)), 1e-300)
        err = np.max(np.abs(a - b)) / scale
        print(f"{what} {f}: max error {err:.3e} relative to max |x|", flush=True)
        assert err <= rtol, f"{what} {f}: max error {err:.3e} (relative to max |x|) > {rtol}"
    assert np.array_equal(g.bond_intact(), o.bond_intact()), f"{what} bond status differs"
    assert np.array_equal(g.contacts(), o.contacts()), f"{what} contact pairs differ"
    wg, rg = g.wall()
    wo, ro = o.wall()
    assert np.array_equal(wg, wo), f"{what} wall positions differ"
    scale = max(np.max(np.abs(ro)), 1e-300)
    assert np.max(np.abs(rg - ro)) / scale <= rtol, f"{what} wall reaction {rg} vs {ro}"


def test_device_present():
    assert _ffi.device_count() >= 1


def test_division_and_sqrt_round_like_host():
    rng = np.random.default_rng(7)
    n = 1 << 20
    a = np.concatenate([rng.uniform(1e-12, 1e-6, n), 10.0 ** rng.uniform(-30, 30, n),
                        rng.uniform(0.5, 2.0, n)])
    b = np.concatenate([rng.uniform(1e-6, 1e-3, n), 10.0 ** rng.uniform(-30, 30, n),
                        rng.uniform(0.5, 2.0, n)])
    q = np.empty_like(a)
    r = np.empty_like(a)
    dp = lambda x: x.ctypes.data_as(_ffi.ctypes.POINTER(_ffi.ctypes.c_double))
    _ffi.check(_ffi.lib().pg_check_arith(0, len(a), dp(a), dp(b), dp(q), dp(r)))
    assert np.array_equal(q, a / b)
    assert np.array_equal(r, np.sqrt(a))


def _red(a, b):
    if a.shape[1] == 2:
        return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def _bond_host(xi, eta):
    """The reference's bond arithmetic (particle2.h:134-146) with the host's
    IEEE operators: s = (|eta| - |xi|)/|xi|, e = eta/|eta|."""
    with np.errstate(all="ignore"):
        xin = np.sqrt(_red(xi, xi))
        n = np.sqrt(_red(eta, eta))
        return (n - xin) / xin, eta / n[:, None]


def _bond_device(dim, check, xi, eta):
    n = len(xi)
    s = np.empty(n)
    e = np.empty((n, dim))
    dp = lambda x: x.ctypes.data_as(_ffi.ctypes.POINTER(_ffi.ctypes.c_double))
    xi, eta = np.ascontiguousarray(xi), np.ascontiguousarray(eta)
    _ffi.check(_ffi.lib().pg_check_bond(0, dim, check, n, dp(xi), dp(eta), dp(s), dp(e)))
    return s, e


def _same(a, b):
    # equal values (a zero quotient's sign may differ, which no force or
    # break test can see) or NaN on both sides
    return (a == b) | (np.isnan(a) & np.isnan(b))


@pytest.mark.parametrize("dim", [2, 3])
def test_bond_arithmetic_at_its_range_edges(dim):
    """The bond loop's own arithmetic (bond_geom: shortened division and
    square-root sequences behind its range test, IEEE operators outside)
    against the host's IEEE results, bit for bit, on inputs chosen at the
    edges of the range test: |xi|^2 and |eta|^2 around 2^-600 and 2^601,
    components of eta around 2^-300, subnormals, zeros, and wide random
    exponents."""
    rng = np.random.default_rng(17 + dim)
    n = 1 << 18

    def comp(lo, hi, size):
        m = rng.uniform(1.0, 2.0, size) * rng.choice([-1.0, 1.0], size)
        return np.ldexp(m, rng.integers(lo, hi + 1, size))

    sets = []
    for lo, hi in ((-303, -297), (297, 303), (-560, 560), (-1074, -1000), (-320, -280),
                   (280, 320)):
        sets.append((comp(lo, hi, (n, dim)), comp(lo, hi, (n, dim))))
    # big and tiny components mixed within one vector, and exact zeros
    a = comp(-305, 305, (n, dim))
    b = comp(-305, 305, (n, dim))
    a[rng.uniform(size=a.shape) < 0.2] = 0.0
    b[rng.uniform(size=b.shape) < 0.2] = 0.0
    sets.append((a, b))
    xi = np.concatenate([x for x, _ in sets])
    eta = np.concatenate([y for _, y in sets])
    s_d, e_d = _bond_device(dim, 1, xi, eta)
    s_h, e_h = _bond_host(xi, eta)
    assert np.all(_same(s_d, s_h)), f"stretch differs at {np.argwhere(~_same(s_d, s_h))[:3].ravel()}"
    assert np.all(_same(e_d, e_h)), "direction differs"
    # the shortcut taken when every coordinate is nice (zero or of magnitude
    # in [2^-240, 2^240]): xi and eta as the kernel forms them from nice X, u
    X = comp(-240, 239, (2, n, dim))
    U = comp(-240, 239, (2, n, dim))
    X[rng.uniform(size=X.shape) < 0.1] = 0.0
    U[rng.uniform(size=U.shape) < 0.3] = 0.0
    for k in range(6):  # the extremes
        X[k % 2, k::97, k % dim] = np.ldexp(1.0, 239 - 478 * (k % 2)) * (1 + 0.1 * k)
    xi2 = X[1] - X[0]
    eta2 = (U[1] - U[0]) + xi2
    keep = np.any(xi2 != 0, axis=1)
    xi2, eta2 = xi2[keep], eta2[keep]
    eta2[::1009] = 0.0  # coincident nodes: s = -1, e NaN
    s_d, e_d = _bond_device(dim, 0, xi2, eta2)
    s_h, e_h = _bond_host(xi2, eta2)
    assert np.all(s_d[::1009] == -1.0) and np.all(np.isnan(e_d[::1009]))
    assert np.all(_same(s_d, s_h)) and np.all(_same(e_d, e_h))


def test_two_disk_collision_bit_exact(oracle_threads):
    s, p = setups.two_disk_collision(seed=3, gap=5e-5)
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (1, 50, 200, 400):
        _advance(g, o, t - done)
        done = t
        _assert_exact(g, o, f"t={t}")
    assert len(o.contacts()) > 0, "workload should be in contact by t=400"


def test_reference_two_disk_case_full_run_bit_exact(oracle_threads):
    # the reference's own small case (2 disks at +-10 m/s, dt 1e-8) over its
    # whole 5000-step run
    s, p = setups.two_disk_collision(seed=0)
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (1000, 2500, 5000):
        _advance(g, o, t - done)
        done = t
        _assert_exact(g, o, f"t={t}")


def test_two_disk_fracture_bit_exact(oracle_threads):
    s, p = setups.two_disk_collision(seed=4, gap=1e-5, speed=50.0)
    # disk 0 expands radially: its bonds stretch past s0 within ~30 steps
    nodes = s.grain_nodes(0)
    s.vel[nodes] += 1e5 * (s.pos[nodes] - s.pos[nodes].mean(axis=0))
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (20, 100, 250, 1000):
        _advance(g, o, t - done)
        done = t
        _assert_exact(g, o, f"t={t}")
    assert (o.bond_intact() == 0).sum() > 0, "workload should break bonds"


@pytest.mark.parametrize("lds", ["1", "0"])
def test_3d_compression_with_walls_bit_exact(oracle_threads, monkeypatch, lds):
    # lds "0" runs the step kernel that gathers partner records from global
    # memory instead of from a workgroup's LDS window
    monkeypatch.setenv("PG_LDS", lds)
    s, p = setups.compression_3d(seed=5, n_side=2, R=0.5e-3, h_div=5, cube=3e-3, wall_speed=-5.0)
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (1, 60, 150):
        _advance(g, o, t - done)
        done = t
        _assert_exact(g, o, f"t={t}")
    _, r = o.wall()
    assert np.any(r != 0), "top wall should push"


def test_3d_compression_long_horizon_bit_exact(oracle_threads):
    # 2000 steps of 8 spheres under a top wall moving at 20 m/s: contact
    # lists change as the pack is pressed together
    s, p = setups.compression_3d(seed=5, n_side=2, R=0.5e-3, h_div=5, cube=3e-3, wall_speed=-20.0)
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (500, 1000, 2000):
        _advance(g, o, t - done)
        done = t
        _assert_exact(g, o, f"t={t}")
    _, r = o.wall()
    assert r[5, 2] > 0, "top wall should push"


def test_3d_wall_reaction_per_call_across_list_rebuilds(oracle_threads):
    # pg_step queues its last step's wall reaction before it learns whether
    # the batch stopped early on a stale neighbour list; then the redone
    # steps must queue it again.  Fast grains and a small Verlet skin force
    # rebuilds inside short calls; after every call the reaction (and the
    # state) must equal the oracle's at that step.
    s, p = setups.compression_3d(seed=5, n_side=2, R=0.5e-3, h_div=5, cube=3e-3, wall_speed=-20.0)
    rng = np.random.default_rng(3)
    for gi in range(s.n_grains):
        s.vel[s.grain_nodes(gi)] = rng.normal(size=3) * 20.0
    g, o = _both(s, p, oracle_threads)
    g.set_tuning(0.25, 256)
    redone_calls = 0
    for n in (1, 1, 3, 7, 17, 40, 29, 64, 65, 33, 50):
        before = g.stats()["skipped_steps"]
        _advance(g, o, n)
        redone_calls += g.stats()["skipped_steps"] > before
        _assert_exact(g, o, f"t={g.t}")
    assert g.stats()["rebuilds"] > 3, "the grains should force list rebuilds"
    assert redone_calls > 0, "some call should stop on a stale list and redo its steps"


@pytest.mark.parametrize("n_grains,box", [(100, (20e-3, 40e-3))])
def test_two_lanes_per_node_mid_size_2d(oracle_threads, n_grains, box):
    # 30k nodes: the step kernel gives each node two lanes (16k < N <= 64k),
    # which share its bond loop; grains kicked to 50 m/s so that contacts,
    # walls and list rebuilds all happen within the horizon
    s, p = setups.bulk_settle_2d(seed=2, n_grains=n_grains, box=box)
    assert 16384 < s.n_nodes <= 65536
    rng = np.random.default_rng(4)
    for gi in range(s.n_grains):
        a = rng.uniform(0, 2 * np.pi)
        s.vel[s.grain_nodes(gi)] = 50.0 * np.array([np.cos(a), np.sin(a)])
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (1, 100, 300):
        _advance(g, o, t - done)
        done = t
        _assert_close(g, o, 1e-10, f"t={t}")
    assert len(o.contacts()) > 0, "grains should touch"


def test_3d_batched_wall_reaction_long_chains_equal_single_runs():
    # three 3D compressions in one handle: each instance's top-wall list in
    # the reaction kernel's piece is a range of the shared side list, long
    # enough to be summed by a wavefront; every instance's reaction (and
    # state) must equal its own single-instance run bit for bit
    sp = [setups.compression_3d(seed=sd, n_side=2, R=0.5e-3, h_div=5, cube=3e-3, wall_speed=-20.0)
          for sd in (5, 6, 7)]
    ss = [s for s, _ in sp]
    ps = [p for _, p in sp]
    steps = 400
    with Stepper(ss, ps, device=0) as b:
        b.step(steps)
        long_lists = 0
        for k, (s, p) in enumerate(sp):
            x = b.read("CurrPos", inst=k)
            wpos, _ = b.wall(inst=k)
            hits = int((np.abs(x[:, 2] - wpos[5]) <= s.contact_radius).sum())
            long_lists += hits > 32  # direction chain of more than 64 terms
            with Stepper(s, p, device=0) as one:
                one.step(steps)
                for f in FIELDS:
                    assert np.array_equal(b.read(f, inst=k), one.read(f)), f"instance {k} {f}"
                w1, r1 = one.wall()
                wb, rb = b.wall(inst=k)
                assert np.array_equal(wb, w1), f"instance {k} wall positions"
                assert np.array_equal(rb, r1), f"instance {k} wall reaction {rb[5]} vs {r1[5]}"
                assert rb[5, 2] != 0, f"instance {k}: the top wall should push"
        assert long_lists >= 2, "the top walls should hold long lists"


def test_3d_batched_instances_against_the_oracle(oracle_threads):
    """Three 3D compressions of different seeds, wall speeds and sizes in one
    handle, one with a sphere expanding so that its bonds break: every
    instance against the oracle's run of it, bit for bit (fields, bond
    status, contact pairs, wall positions and reaction) at three steps."""
    sp = [setups.compression_3d(seed=11, n_side=2, R=0.5e-3, h_div=5, cube=3e-3, wall_speed=-20.0),
          setups.compression_3d(seed=12, n_side=2, R=0.6e-3, h_div=5, cube=3.5e-3, wall_speed=-5.0),
          setups.compression_3d(seed=13, n_side=3, R=0.4e-3, h_div=4, cube=4e-3, wall_speed=-30.0)]
    s1 = sp[1][0]
    nodes = s1.grain_nodes(2)
    s1.vel[nodes] += 6e4 * (s1.pos[nodes] - s1.pos[nodes].mean(axis=0))
    ss = [s for s, _ in sp]
    ps = [p for _, p in sp]
    print(f"instances: {[s.n_nodes for s in ss]} nodes", flush=True)
    os_ = [oracle.Oracle(s, p, threads=oracle_threads) for s, p in sp]
    try:
        with Stepper(ss, ps, device=0) as b:
            done = 0
            for t in (1, 60, 150):
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
                    wb, rb = b.wall(inst=k)
                    wo, ro = o.wall()
                    assert np.array_equal(wb, wo), f"instance {k} t={t} wall positions"
                    assert np.array_equal(rb, ro), f"instance {k} t={t} wall reaction"
                done = t
            assert int((os_[1].bond_intact() == 0).sum()) > 100, "instance 1 should break bonds"
            assert any(np.any(o.wall()[1] != 0) for o in os_), "a wall should push"
    finally:
        for o in os_:
            o.close()


def _falling_disk(seed, reset_step):
    # disk 0 starts 4R above disk 1 and falls at 50 m/s; at reset_step
    # reset_partzero_y moves it (pos shifted in place, bonds unchanged) to
    # sit 0.9R + R_c above disk 1's top, so it lands on disk 1
    s, p = setups.two_disk_collision(seed=seed, gap=5e-5)
    a = int(s.node_offset[1])
    s.pos[:a] += np.array([s.pos[a:, 0].mean() - s.pos[:a, 0].mean(), 4e-3])
    s.vel[:] = 0.0
    s.vel[:a] = [0.0, -50.0]
    p.reset_partzero_y, p.reset_partzero_y_timestep, p.wheel_rad = 1, reset_step, 0.9e-3
    return s, p


def test_reset_partzero_y_bit_exact(oracle_threads):
    s, p = _falling_disk(3, 30)
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (29, 30, 31, 400):
        _advance(g, o, t - done)
        done = t
        _assert_exact(g, o, f"t={t}")
    assert len(o.contacts()) > 0, "the moved disk should touch the other"


def test_reset_partzero_y_batched_equals_single():
    sp = [_falling_disk(3, 30), _falling_disk(4, 45)]
    ss = [s for s, _ in sp]
    ps = [p for _, p in sp]
    with Stepper(ss, ps, device=0) as b:
        b.step(300)
        for k, (s, p) in enumerate(sp):
            with Stepper(s, p, device=0) as one:
                one.step(25)
                one.step(275)
                for f in FIELDS:
                    assert np.array_equal(b.read(f, inst=k), one.read(f)), f"instance {k} {f}"
                assert np.array_equal(b.contacts(inst=k), one.contacts())


def test_worm_loop_bit_exact(oracle_threads):
    # run_timeloop_worm across the phase change at step 1000 (front nodes
    # held and extforce on before it, back nodes held and extforce off after)
    s, p = setups.two_disk_collision(seed=3, gap=5e-5)
    s.extforce[:] = [0.0, -9.81 * 1200.0]
    p.timeloop_variant = 1
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (10, 999, 1000, 1001, 1100):
        _advance(g, o, t - done)
        done = t
        _assert_exact(g, o, f"t={t}")
    assert len(o.contacts()) > 0, "the disks should be in contact"


def test_compact_loop_bit_exact(oracle_threads):
    # run_timeloop_compact: disk 1 is not stoppable but still makes contact,
    # the extforce ramp acts, torque does not exist in this loop
    s, p = setups.two_disk_collision(seed=3, gap=5e-5)
    s.extforce[:] = [0.0, -9.81 * 1200.0]
    s.stoppable[1] = 0
    s.torque_axis = np.full(s.n_grains, 2, dtype=np.uint32)
    s.torque_val = np.full(s.n_grains, 1e6)
    p.timeloop_variant, p.gradient_extforce, p.extforce_maxstep, p.enable_torque = 2, 1, 200, 1
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (1, 100, 400):
        _advance(g, o, t - done)
        done = t
        _assert_exact(g, o, f"t={t}")
    assert len(o.contacts()) > 0, "the disks should be in contact"


def test_2d_settle_with_walls_close(oracle_threads):
    s, p = setups.bulk_settle_2d(seed=6, n_grains=16, box=(8e-3, 16e-3), nodes_per_grain=80)
    # start the grains at 100 m/s in random directions so that they collide
    # with each other and with the walls within a few hundred steps
    rng = np.random.default_rng(0)
    for gi in range(s.n_grains):
        a = rng.uniform(0, 2 * np.pi)
        s.vel[s.grain_nodes(gi)] = 100.0 * np.array([np.cos(a), np.sin(a)])
    g, o = _both(s, p, oracle_threads)
    done = 0
    for t in (1, 200, 600, 1000):
        _advance(g, o, t - done)
        done = t
        _assert_close(g, o, 1e-10, f"t={t}")


def test_batched_equals_single():
    sp = [setups.settle_instance_2d(seed) for seed in (11, 12, 13)]
    ss = [setups.subset(s, list(range(min(12, s.n_grains)))) for s, _ in sp]
    ps = [p for _, p in sp]
    steps = 120
    with Stepper(ss, ps, device=0) as b:
        b.step(steps)
        for k, (s, p) in enumerate(zip(ss, ps)):
            with Stepper(s, p, device=0) as one:
                one.step(steps)
                for f in FIELDS:
                    assert np.array_equal(b.read(f, inst=k), one.read(f)), f"instance {k} {f}"
                assert np.array_equal(b.bond_intact(inst=k), one.bond_intact())
                assert np.array_equal(b.contacts(inst=k), one.contacts())
                wb, rb = b.wall(inst=k)
                w1, r1 = one.wall()
                assert np.array_equal(wb, w1) and np.array_equal(rb, r1)
