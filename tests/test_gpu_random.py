"""Randomised GPU parity: seeded draws of a workload, grain velocities, an
optional expanding grain (bond breaks), and up to three of the configuration
switches of test_gpu_options at once, each run against the oracle, alone and
as instances of batched handles.  The
workloads span every step-kernel form: small handles (four lanes per node),
a 2D handle of about 30k nodes (two lanes) and one of about 75k (one lane per
node).  Bit for bit, except 1e-10 of each field's largest value for 2D
workloads with walls (the wall force calls libm asin, whose last bit may
differ between host and device); bond status, contact pairs and wall
positions exact in every case."""
import os

import numpy as np
import pytest

import oracle
import test_gpu_options as opt
from perigrain_amd import Stepper, setups
from test_gpu_parity import FIELDS, _advance, _assert_close, _assert_exact, _both

pytestmark = pytest.mark.gpu

# PG_RANDOM_CASES widens the sweep (the draws are the same for the first 20)
N_CASES = int(os.environ.get("PG_RANDOM_CASES", "20"))

# switches that apply to every workload, and those for 3D only
MODS = [opt._nl_bdry_only, opt._no_damping_no_friction, opt._no_friction, opt._no_self_contact,
        opt._grain0_not_breakable, opt._timed_toggles, opt._torque, opt._clamped,
        opt._nonuniform_volume, opt._extforce_ramp, opt._bdry_set_order]
MODS_3D = [opt._torque_axes_x_y]


def _draw(case):
    rng = np.random.default_rng(1000 + case)
    kind = ("2d-bulk", "2d-disks", "3d", "2d-mid", "2d-large")[case % 5]
    seed = int(rng.integers(0, 1 << 30))
    if kind == "2d-bulk":
        s, p = setups.bulk_settle_2d(seed=seed, n_grains=int(rng.integers(4, 13)),
                                     box=(8e-3, 16e-3), nodes_per_grain=int(rng.integers(40, 160)))
    elif kind == "2d-disks":
        s, p = setups.two_disk_collision(seed=seed, gap=float(rng.uniform(1e-5, 6e-5)),
                                         speed=float(rng.uniform(10.0, 60.0)))
    elif kind == "3d":
        s, p = setups.compression_3d(seed=seed, n_side=2, R=0.5e-3, h_div=int(rng.integers(4, 6)),
                                     cube=3e-3, wall_speed=-float(rng.uniform(5.0, 30.0)))
    elif kind == "2d-mid":
        s, p = setups.bulk_settle_2d(seed=seed, n_grains=100, box=(25e-3, 50e-3),
                                     nodes_per_grain=300)
    else:
        s, p = setups.bulk_settle_2d(seed=seed, n_grains=250, box=(40e-3, 80e-3),
                                     nodes_per_grain=300)
    # every grain thrown in a random direction, so that contacts start early
    for g in range(s.n_grains):
        s.vel[s.grain_nodes(g)] += rng.normal(size=s.dim) * float(rng.uniform(20.0, 150.0))
    names = []
    if rng.uniform() < 0.5:  # one grain expands until its bonds break
        g = int(rng.integers(0, s.n_grains))
        nodes = s.grain_nodes(g)
        s.vel[nodes] += float(rng.uniform(3e4, 1e5)) * (s.pos[nodes] - s.pos[nodes].mean(axis=0))
        names.append(f"expand-{g}")
    pool = MODS + (MODS_3D if s.dim == 3 else [])
    for k in rng.choice(len(pool), size=int(rng.integers(0, 4)), replace=False):
        pool[k](s, p)
        names.append(pool[k].__name__.strip("_"))
    # one case in four follows run_timeloop_compact or run_timeloop_worm
    # (drawn from a stream of its own, so the draws above stay as they were)
    loop = int(np.random.default_rng(2000 + case).choice([0, 0, 0, 0, 0, 0, 2, 1]))
    if loop and np.all(np.diff(s.node_offset) >= 4):
        p.timeloop_variant = loop
        names.append(("", "worm-loop", "compact-loop")[loop])
    return kind, s, p, names


def _draw_small(rng, dim):
    """A small workload of the given dimension with random velocities, maybe
    an expanding grain, and 0-3 configuration switches (as _draw)."""
    seed = int(rng.integers(0, 1 << 30))
    if dim == 3:
        s, p = setups.compression_3d(seed=seed, n_side=2, R=0.5e-3, h_div=int(rng.integers(4, 6)),
                                     cube=3e-3, wall_speed=-float(rng.uniform(5.0, 30.0)))
    elif rng.uniform() < 0.5:
        s, p = setups.bulk_settle_2d(seed=seed, n_grains=int(rng.integers(4, 13)),
                                     box=(8e-3, 16e-3), nodes_per_grain=int(rng.integers(40, 160)))
    else:
        s, p = setups.two_disk_collision(seed=seed, gap=float(rng.uniform(1e-5, 6e-5)),
                                         speed=float(rng.uniform(10.0, 60.0)))
    for g in range(s.n_grains):
        s.vel[s.grain_nodes(g)] += rng.normal(size=s.dim) * float(rng.uniform(20.0, 150.0))
    names = []
    if rng.uniform() < 0.5:
        g = int(rng.integers(0, s.n_grains))
        nodes = s.grain_nodes(g)
        s.vel[nodes] += float(rng.uniform(3e4, 1e5)) * (s.pos[nodes] - s.pos[nodes].mean(axis=0))
        names.append(f"expand-{g}")
    pool = MODS + (MODS_3D if dim == 3 else [])
    for k in rng.choice(len(pool), size=int(rng.integers(0, 4)), replace=False):
        pool[k](s, p)
        names.append(pool[k].__name__.strip("_"))
    return s, p, names


@pytest.mark.parametrize("case", range(6))
def test_random_batched_instances_match_the_oracle(case, oracle_threads):
    """Three to five random small instances of one dimension, each with its
    own switches, time step and walls, in one batched handle: every instance
    against the oracle's run of it (bit for bit; 1e-10 for 2D instances with
    walls)."""
    rng = np.random.default_rng(5000 + case)
    dim = 3 if case % 3 == 2 else 2
    draws = [_draw_small(rng, dim) for _ in range(int(rng.integers(3, 6)))]
    what = f"batch {case}: " + ", ".join(f"{s.n_nodes} nodes {'+'.join(n) or 'plain'}"
                                          for s, _, n in draws)
    print(what, flush=True)
    os_ = [oracle.Oracle(s, p, threads=oracle_threads) for s, p, _ in draws]
    try:
        with Stepper([s for s, _, _ in draws], [p for _, p, _ in draws], device=0) as b:
            done = 0
            for t in (1, 60, 150):
                b.step(t - done)
                for k, (o, (s, _, _)) in enumerate(zip(os_, draws)):
                    o.step(t - done, record_contacts=True)
                    close = s.dim == 2 and s.allow_wall
                    for f in FIELDS:
                        a, c = b.read(f, inst=k), o.read(f)
                        if close:
                            err = np.max(np.abs(a - c)) / max(np.max(np.abs(c)), 1e-300)
                            assert err <= 1e-10, f"instance {k} t={t} {f}: {err:.3e}"
                        else:
                            assert np.array_equal(a, c), \
                                f"instance {k} t={t} {f}: {int((a != c).sum())} entries differ"
                    assert np.array_equal(b.bond_intact(inst=k), o.bond_intact()), \
                        f"instance {k} t={t} bond status"
                    assert np.array_equal(b.contacts(inst=k), o.contacts()), \
                        f"instance {k} t={t} contact pairs"
                    wb, rb = b.wall(inst=k)
                    wo, ro = o.wall()
                    assert np.array_equal(wb, wo), f"instance {k} t={t} wall positions"
                    if close:
                        scale = max(np.max(np.abs(ro)), 1e-300)
                        assert np.max(np.abs(rb - ro)) / scale <= 1e-10, f"instance {k} t={t} reaction"
                    else:
                        assert np.array_equal(rb, ro), f"instance {k} t={t} wall reaction"
                done = t
        print(f"batch {case}: " + ", ".join(
            f"{len(o.contacts())} pairs/{int((o.bond_intact() == 0).sum())} broken" for o in os_),
            flush=True)
    finally:
        for o in os_:
            o.close()


@pytest.mark.parametrize("case", range(N_CASES))
def test_random_configuration_matches_the_oracle(case, oracle_threads):
    kind, s, p, names = _draw(case)
    what = f"case {case} ({kind}, {s.n_nodes} nodes, {s.n_grains} grains, {'+'.join(names) or 'plain'})"
    print(what, flush=True)
    g, o = _both(s, p, oracle_threads)
    try:
        done = 0
        for t in (1, 60, 150):
            _advance(g, o, t - done)
            done = t
            if s.dim == 2 and s.allow_wall:
                _assert_close(g, o, 1e-10, f"{what} t={t}")
            else:
                _assert_exact(g, o, f"{what} t={t}")
        print(f"{what}: {len(o.contacts())} contact pairs, "
              f"{int((o.bond_intact() == 0).sum())} broken slots at t=150", flush=True)
    finally:
        g.close()
        o.close()
