"""Parity at the benchmark sizes.

* 3D compression of 64 spheres (264,896 nodes, 23.5M directed bonds), with
  one sphere expanding so that bonds break from the first step on: the first
  steps bit for bit against the oracle (every field, bond status, contact
  pairs, walls), then properties that hold at any size over a longer GPU run
  -- the broken-bond set is symmetric (s_ij == s_ji exactly) and only grows,
  each grain's Connectivity is its intact bonds in adjacency order, the
  state stays finite.
* The batched form at 256 instances (2.29M nodes), graph-replayed and
  autotuned: sampled instances bit for bit against their own
  single-instance runs.
* The differentiable form on the 2000-step 2D compression (64 grains):
  the adjoint gradient against the FP64 CPU dual-number gradient (all three
  parameters) and against central differences (K and the stiffness scale;
  the friction cap's jump at its threshold makes central differences in mu
  inexact, so mu is held to the CPU gradient).
"""
import numpy as np
import pytest

import oracle
from perigrain_amd import Stepper, diffrun, setups

pytestmark = pytest.mark.gpu

FIELDS = ("CurrPos", "vel", "acc", "force")


def directed_slots(s):
    """(source, partner) node of every directed bond slot, instance-global,
    in the order of the bond-status readers: grain, node, then position in
    the node's row, where each Connectivity row (i, j) appends j to row i
    and i to row j (conn2NArr, load_hdf5.h:28-38)."""
    src, dst = [], []
    for g in range(s.n_grains):
        rows = s.conn[s.conn_offset[g]:s.conn_offset[g + 1]].astype(np.int64) + s.node_offset[g]
        a = np.empty(2 * len(rows), np.int64)
        b = np.empty_like(a)
        a[0::2], b[0::2] = rows[:, 0], rows[:, 1]
        a[1::2], b[1::2] = rows[:, 1], rows[:, 0]
        order = np.argsort(a, kind="stable")
        src.append(a[order])
        dst.append(b[order])
    return np.concatenate(src), np.concatenate(dst)


def mirror_slots(src, dst, n):
    """Index of the slot (j, i) for every slot (i, j)."""
    key = src * n + dst
    order = np.argsort(key)
    pos = np.searchsorted(key[order], dst * n + src)
    m = order[pos]
    assert np.array_equal(key[m], dst * n + src), "every bond has a reverse slot"
    return m


def expected_conn(s, src, dst, intact, g):
    """Connectivity of grain g as NbdArr2conn writes it (timeloop.h:25-49):
    intact slots (i, q) with i < q in slot order, grain-local."""
    lo, hi = 2 * int(s.conn_offset[g]), 2 * int(s.conn_offset[g + 1])
    keep = (intact[lo:hi] == 1) & (src[lo:hi] < dst[lo:hi])
    off = int(s.node_offset[g])
    return np.stack([src[lo:hi][keep] - off, dst[lo:hi][keep] - off], axis=1).astype(np.uint32)


def _expanding_compression():
    # the benchmark workload; sphere 0 expands radially at 6e5 1/s with 50%
    # per-node spread, so its bond stretches cross s0 over steps 1-4
    s, p = setups.compression_3d(seed=0)
    rng = np.random.default_rng(11)
    nd = s.grain_nodes(0)
    r = s.pos[nd] - s.pos[nd].mean(axis=0)
    s.vel[nd] += 6e5 * r * (1.0 + 0.5 * rng.standard_normal((r.shape[0], 1)))
    return s, p


def test_compression_3d_full_size(oracle_threads):
    s, p = _expanding_compression()
    print(f"3D compression, {s.n_nodes} nodes, {s.n_slots()} directed bonds", flush=True)
    with Stepper(s, p, device=0) as g:
        o = oracle.Oracle(s, p, threads=oracle_threads)
        try:
            done = 0
            for t in (1, 3):
                g.step(t - done)
                o.step(t - done, record_contacts=True)
                done = t
                for f in FIELDS:
                    a, b = g.read(f), o.read(f)
                    bad = np.argwhere(a != b)
                    assert len(bad) == 0, f"t={t} {f}: {len(bad)} entries differ, first " \
                                          f"{bad[:3].tolist()}"
                bi = g.bond_intact()
                assert np.array_equal(bi, o.bond_intact()), f"t={t} bond status differs"
                assert np.array_equal(g.contacts(), o.contacts()), f"t={t} contact pairs differ"
                wg, rg = g.wall()
                wo, ro = o.wall()
                assert np.array_equal(wg, wo) and np.array_equal(rg, ro), f"t={t} walls differ"
            assert 0 < int((bi == 0).sum()) < 2 * int(s.conn_offset[1]), \
                "sphere 0 should have broken some, not all, of its bonds by t=3"
            assert len(o.contacts()) > 0, "neighbouring spheres start within contact range"
        finally:
            o.close()

        src, dst = directed_slots(s)
        mirror = mirror_slots(src, dst, s.n_nodes)
        before = bi
        g.step(197)
        bi = g.bond_intact()
        assert np.array_equal(bi, bi[mirror]), "broken-bond set is not symmetric"
        assert np.all(bi <= before), "a broken bond came back"
        assert int((bi == 0).sum()) > int((before == 0).sum()), "breaking should go on"
        for gr in (0, 1, s.n_grains - 1):
            assert np.array_equal(g.conn(gr), expected_conn(s, src, dst, bi, gr)), \
                f"Connectivity of grain {gr}"
        for f in FIELDS:
            assert np.all(np.isfinite(g.read(f))), f"{f} not finite at t=200"
        assert g.t == 200


def _assert_instance(b, k, o, what):
    """Instance k of a handle against an oracle run of that instance: bond
    status, contact pairs and wall positions bit for bit; the fields and the
    wall reaction bit for bit unless a node touched a 2D wall (asin, whose
    last bit the host and device libraries may round differently), then to
    1e-10 of each field's largest value."""
    walls = bool(np.any(o.wall()[1] != 0))
    for f in FIELDS:
        a, c = b.read(f, inst=k), o.read(f)
        if walls:
            err = np.max(np.abs(a - c)) / max(np.max(np.abs(c)), 1e-300)
            assert err <= 1e-10, f"{what} {f}: {err:.3e}"
        else:
            assert np.array_equal(a, c), f"{what} {f}: {int((a != c).sum())} entries differ"
    assert np.array_equal(b.bond_intact(inst=k), o.bond_intact()), f"{what} bond status"
    assert np.array_equal(b.contacts(inst=k), o.contacts()), f"{what} contact pairs"
    wb, rb = b.wall(inst=k)
    wo, ro = o.wall()
    assert np.array_equal(wb, wo), f"{what} wall positions"
    if walls:
        assert np.max(np.abs(rb - ro)) <= 1e-10 * np.max(np.abs(ro)), f"{what} wall reaction"
    else:
        assert np.array_equal(rb, ro), f"{what} wall reaction"
    return walls


def test_batched_256_instances_equal_single_runs(oracle_threads):
    S, P = setups.batched_settle(256, seed=0)
    sizes = [x.n_nodes for x in S]
    small, large = int(np.argmin(sizes)), int(np.argmax(sizes))
    print(f"batched: {len(S)} instances, {sum(sizes)} nodes; smallest {small} "
          f"({sizes[small]} nodes), largest {large} ({sizes[large]} nodes)", flush=True)
    with Stepper(S, P, device=0) as b:
        b.set_graph(True)
        b.autotune(trial_steps=8)  # advances the run by 9 * 8 steps
        b.step(160 - b.t)
        assert b.t == 160
        # instances of the batched handle against the oracle's run of each
        # (timeloop.h:608-878 per instance), bit for bit
        for k in sorted({small, large, 0, 97, 255}):
            o = oracle.Oracle(S[k], P[k], threads=oracle_threads)
            try:
                o.step(160, record_contacts=True)
                walls = _assert_instance(b, k, o, f"instance {k} vs oracle")
                print(f"instance {k}: {sizes[k]} nodes, {len(o.contacts())} contact pairs, "
                      f"{'wall contact' if walls else 'no wall contact'}: matches the oracle",
                      flush=True)
            finally:
                o.close()
        for k in (0, 1, 97, 200, 255):
            with Stepper(S[k], P[k], device=0) as one:
                one.step(160)
                for f in FIELDS:
                    assert np.array_equal(b.read(f, inst=k), one.read(f)), f"instance {k} {f}"
                assert np.array_equal(b.bond_intact(inst=k), one.bond_intact()), f"instance {k}"
                assert np.array_equal(b.contacts(inst=k), one.contacts()), f"instance {k}"
                wb, rb = b.wall(inst=k)
                w1, r1 = one.wall()
                assert np.array_equal(wb, w1) and np.array_equal(rb, r1), f"instance {k} walls"


def _settle_2d_against_oracle(s, p, checkpoints, threads, what):
    print(f"{what}: {s.n_nodes} nodes, {s.n_grains} grains, {s.n_slots()} directed bonds",
          flush=True)
    from test_gpu_parity import _assert_close
    with Stepper(s, p, device=0) as g:
        o = oracle.Oracle(s, p, threads=threads)
        try:
            done = 0
            for t in checkpoints:
                g.step(t - done)
                o.step(t - done, record_contacts=True)
                done = t
                exact = all(np.array_equal(g.read(f), o.read(f)) for f in FIELDS)
                print(f"{what} t={t}: fields {'bit-identical' if exact else 'within tolerance'}, "
                      f"{len(o.contacts())} contact pairs, "
                      f"{int((o.bond_intact() == 0).sum())} broken slots", flush=True)
                _assert_close(g, o, 1e-10, f"{what} t={t}")
            return o.contacts(), o.bond_intact(), o.wall()[1]
        finally:
            o.close()


def test_bulk_settle_2d_full_size(oracle_threads):
    # the 2D bulk settle at its own size and settings: 400 grains (disks,
    # squares, plus shapes, annuli), 121,900 nodes, gravity, 4 walls,
    # fracture on, dt 2e-8, 1000 steps against the oracle
    s, p = setups.bulk_settle_2d(seed=0)
    assert s.n_grains == 400 and p.enable_fracture == 1 and s.allow_wall == 1
    _settle_2d_against_oracle(s, p, (1, 500, 1000), oracle_threads, "bulk settle")


def test_bulk_settle_2d_full_size_colliding(oracle_threads):
    # the same workload with every grain thrown at 20 m/s in a random
    # direction, so that grains collide with each other and with the walls
    # within the horizon
    s, p = setups.bulk_settle_2d(seed=0)
    rng = np.random.default_rng(1)
    for gi in range(s.n_grains):
        a = rng.uniform(0, 2 * np.pi)
        s.vel[s.grain_nodes(gi)] = 20.0 * np.array([np.cos(a), np.sin(a)])
    pairs, _, react = _settle_2d_against_oracle(s, p, (1, 300, 600), oracle_threads,
                                                "bulk settle, colliding")
    assert len(pairs) > 0, "grains should collide"
    assert np.any(react != 0), "grains should reach the walls"


def test_adjoint_2000_steps(oracle_threads):
    theta0, theta_t = (1.5e9, 1.0, 0.25), (1.0e9, 1.2, 0.4)
    run = diffrun.compression_2d_run(steps=2000, checkpoint_every=100)
    target, st = run.forward(*theta_t)
    st[0].close()
    hist, state = run.forward(*theta0)
    _, bar = diffrun.top_wall_misfit(hist, target)
    grad, raw = run.backward(state, bar)

    s, p = run.make_setup(*theta0)
    R, dR = oracle.reaction_history_tangent(s, p, run.steps, threads=oracle_threads)
    np.testing.assert_allclose(hist, R, rtol=1e-10, atol=1e-10 * np.abs(R).max())
    bar_c = np.zeros_like(R)
    bar_c[:, 2] = 2.0 * (R[:, 2] - target[:, 2])
    raw_c = np.einsum("tsk,tskq->q", bar_c, dR)
    print("adjoint raw", raw, "cpu dual numbers", raw_c, flush=True)
    np.testing.assert_allclose(raw, raw_c, rtol=1e-8)

    for k in (0, 1):
        h = 1e-6 * theta0[k]
        vals = []
        for sgn in (1, -1):
            th = list(theta0)
            th[k] += sgn * h
            hh, stt = run.forward(*th)
            stt[0].close()
            vals.append(diffrun.top_wall_misfit(hh, target)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"parameter {k}: adjoint {grad[k]!r} central differences {fd!r}", flush=True)
        assert abs(grad[k] - fd) <= 1e-6 * abs(fd)
