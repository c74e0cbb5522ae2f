"""Differentiable form on a GPU: the recorded forward pass reproduces the
plain run, and the adjoint gradient of a wall-reaction misfit agrees with
central finite differences (relative step 1e-6)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")  # before libpg: one HIP runtime for both

from perigrain_amd import adjoint, diffrun  # noqa: E402

pytestmark = pytest.mark.gpu

THETA0 = (1.5e9, 1.0, 0.25)
THETA_T = (1.0e9, 1.2, 0.4)


def _small_run(steps=240, ck=50, **kw):
    kw.setdefault("n_grains", 4)
    kw.setdefault("nodes_per_grain", 60)
    return diffrun.compression_2d_run(steps=steps, checkpoint_every=ck, **kw)


def test_recorded_history_equals_plain_steps():
    run = _small_run()
    hist, state = run.forward(*THETA0)
    state[0].close()
    plain = run.plain(*THETA0)
    assert np.any(hist[:, 2, 1] != 0), "the top wall should be in contact"
    assert np.array_equal(hist, plain)


def test_backward_leaves_the_final_state():
    run = _small_run(steps=130, ck=40)
    hist, state = run.forward(*THETA0)
    g = state[0]
    x0 = g.read("CurrPos")
    v0 = g.read("vel")
    g.adjoint_backward(np.ones_like(hist))
    assert g.t == 130
    assert np.array_equal(g.read("CurrPos"), x0) and np.array_equal(g.read("vel"), v0)
    # and it keeps stepping like a run that never went backwards
    g.step(20)
    s, p = run.make_setup(*THETA0)
    from perigrain_amd import Stepper
    with Stepper(s, p, device=0) as ref:
        ref.step(150)
        assert np.array_equal(g.read("CurrPos"), ref.read("CurrPos"))
    g.close()


def _loss(run, theta, target):
    hist = run.plain(*theta)
    return float(((hist[:, 2, :] - target[:, 2, :]) ** 2).sum())


@pytest.mark.parametrize("ck", [50, 240])
def test_gradient_matches_central_differences(ck):
    run = _small_run(ck=ck)
    target = torch.from_numpy(run.plain(*THETA_T))
    theta = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    hist = adjoint.reaction_history(theta, run)
    loss = adjoint.top_wall_misfit(hist, target)
    loss.backward()
    grad = theta.grad.numpy()
    tgt = target.numpy()
    fd = np.zeros(3)
    for k in range(3):
        h = 1e-6 * THETA0[k]
        tp = list(THETA0)
        tm = list(THETA0)
        tp[k] += h
        tm[k] -= h
        fd[k] = (_loss(run, tp, tgt) - _loss(run, tm, tgt)) / (tp[k] - tm[k])
    print("adjoint", grad, "central differences", fd, flush=True)
    rel = np.abs(grad - fd) / np.abs(fd)
    assert np.all(rel < 1e-5), f"adjoint {grad} vs central differences {fd}: rel {rel}"


def test_gradient_matches_cpu_dual_number_reference(oracle_threads):
    # The FP64 CPU reference: the oracle on dual numbers, seeded along the
    # same three parameters, differentiates the same frozen-branch map.
    import oracle
    run = _small_run()
    target = run.plain(*THETA_T)
    hist, state = run.forward(*THETA0)
    _, bar = diffrun.top_wall_misfit(hist, target)
    _, raw = run.backward(state, bar)
    s, p = run.make_setup(*THETA0)
    R, dR = oracle.reaction_history_tangent(s, p, run.steps, threads=oracle_threads)
    np.testing.assert_allclose(hist, R, rtol=1e-10, atol=1e-10 * np.abs(R).max())
    bar_c = np.zeros_like(R)
    bar_c[:, 2] = 2.0 * (R[:, 2] - target[:, 2])
    raw_cpu = np.einsum("tsk,tskq->q", bar_c, dR)
    print("adjoint", raw, "cpu dual numbers", raw_cpu, flush=True)
    np.testing.assert_allclose(raw, raw_cpu, rtol=1e-8)


def test_3d_adjoint_matches_cpu_dual_numbers(oracle_threads):
    # 8 spheres pressed by the top wall (z_max, side 5) from the first step;
    # 100 steps in segments of 30 (the last one shorter)
    import oracle
    from perigrain_amd import Stepper, setups
    s, p = setups.compression_3d(seed=5, n_side=2, R=0.5e-3, h_div=5, cube=3e-3, wall_speed=-50.0)
    steps, ck = 100, 30
    R, dR = oracle.reaction_history_tangent(s, p, steps, threads=oracle_threads)
    with Stepper(s, p, device=0) as g:
        hist = g.adjoint_forward(steps, ck)
        assert np.array_equal(hist, R), "the recorded 3D reaction history is bit-exact"
        assert np.all(np.any(hist[:, 5] != 0, axis=1)), "the top wall should push at every step"
        bar = np.zeros_like(hist)
        bar[:, 5] = 2.0 * hist[:, 5]  # misfit against a zero target
        raw = g.adjoint_backward(bar)
    raw_cpu = np.einsum("tsk,tskq->q", bar, dR)
    print("adjoint", raw, "cpu dual numbers", raw_cpu, flush=True)
    np.testing.assert_allclose(raw, raw_cpu, rtol=1e-8)


def _fracturing_compression():
    # the small 2D compression with fracture on; the grain under the top
    # wall expands radially (2e4 1/s, 50% per-node spread), so its bonds
    # break from step 5 on, inside the first and second checkpoint segments
    from perigrain_amd import setups
    s, p = setups.compression_2d_diff(n_grains=4, nodes_per_grain=60)
    p.enable_fracture = 1
    g = int(np.argmax([s.pos[s.grain_nodes(k), 1].max() for k in range(s.n_grains)]))
    nd = s.grain_nodes(g)
    r = s.pos[nd] - s.pos[nd].mean(axis=0)
    rng = np.random.default_rng(5)
    s.vel[nd] += 2e4 * r * (1.0 + 0.5 * rng.standard_normal((r.shape[0], 1)))
    return s, p


def test_adjoint_with_bonds_breaking_mid_segment(oracle_threads):
    # Bond status frozen to the forward pass, with real breaks: the step of
    # each break is recorded, each segment's recompute restores the bonds
    # broken before it, and the adjoint of step t uses a bond only while it
    # was intact at t (the break happens after the step's force used it).
    # The gradient is checked against the CPU dual-number oracle, which
    # breaks the same bonds on their values, and the state after the
    # backward pass is the state after the forward pass.
    import oracle
    from perigrain_amd import Stepper
    s, p = _fracturing_compression()
    steps, ck = 120, 30
    with Stepper(s, p, device=0) as probe:  # where the breaks happen
        broken = []
        for t in range(1, steps + 1):
            probe.step(1)
            broken.append(int((probe.bond_intact() == 0).sum()))
    new = [t for t in range(1, steps + 1) if broken[t - 1] > (broken[t - 2] if t > 1 else 0)]
    print(f"bonds break at steps {new[:12]}..., {broken[-1]} broken slots at t={steps}", flush=True)
    assert any(t % ck not in (0, 1) for t in new), "a break strictly inside a segment"
    assert len(new) >= 5 and broken[-1] < s.n_slots()
    R, dR = oracle.reaction_history_tangent(s, p, steps, threads=oracle_threads)
    with Stepper(s, p, device=0) as g:
        hist = g.adjoint_forward(steps, ck)
        np.testing.assert_allclose(hist, R, rtol=1e-10, atol=1e-10 * np.abs(R).max())
        assert np.any(hist[:, 2, 1] != 0), "the top wall should be in contact"
        x0, v0, b0 = g.read("CurrPos"), g.read("vel"), g.bond_intact()
        assert int((b0 == 0).sum()) == broken[-1]
        bar = np.zeros_like(hist)
        bar[:, 2] = 2.0 * hist[:, 2]  # misfit against a zero target
        raw = g.adjoint_backward(bar)
        assert g.t == steps
        assert np.array_equal(g.read("CurrPos"), x0) and np.array_equal(g.read("vel"), v0)
        assert np.array_equal(g.bond_intact(), b0), "the broken-bond set after the backward pass"
    bar_c = np.zeros_like(R)
    bar_c[:, 2] = 2.0 * R[:, 2]
    raw_cpu = np.einsum("tsk,tskq->q", bar_c, dR)
    print("adjoint", raw, "cpu dual numbers", raw_cpu, flush=True)
    np.testing.assert_allclose(raw, raw_cpu, rtol=1e-8)


def _random_adjoint_case(case):
    """A small 2D or 3D compression with a random wall speed, length and
    checkpoint interval, maybe fracture with the top grain expanding, and up
    to three of the switches the differentiable form supports."""
    import test_gpu_options as opt
    from perigrain_amd import setups
    mods = [opt._no_damping_no_friction, opt._no_friction, opt._no_self_contact,
            opt._nl_bdry_only, opt._nonuniform_volume, opt._bdry_set_order,
            opt._extforce_ramp, opt._grain0_not_breakable]
    rng = np.random.default_rng(7000 + case)
    seed = int(rng.integers(0, 1 << 30))
    if case % 3 == 2:
        s, p = setups.compression_3d(seed=seed, n_side=2, R=0.5e-3, h_div=4, cube=3e-3,
                                     wall_speed=-float(rng.uniform(20.0, 60.0)))
    else:
        s, p = setups.compression_2d_diff(seed=seed, n_grains=int(rng.integers(3, 7)),
                                          nodes_per_grain=int(rng.integers(40, 90)),
                                          wall_speed=-float(rng.uniform(5.0, 20.0)))
    names = []
    if rng.uniform() < 0.5:
        p.enable_fracture = 1
        g = int(np.argmax([s.pos[s.grain_nodes(k), -1].max() for k in range(s.n_grains)]))
        nd = s.grain_nodes(g)
        r = s.pos[nd] - s.pos[nd].mean(axis=0)
        spread = 1.0 + 0.5 * rng.standard_normal((r.shape[0], 1))
        s.vel[nd] += float(rng.uniform(2e4, 5e4)) * r * spread
        names.append(f"fracture-expand-{g}")
    for k in rng.choice(len(mods), size=int(rng.integers(0, 4)), replace=False):
        mods[k](s, p)
        names.append(mods[k].__name__.strip("_"))
    steps, ck = int(rng.integers(60, 140)), int(rng.integers(15, 50))
    return s, p, steps, ck, names, rng


@pytest.mark.parametrize("case", range(6))
def test_random_adjoint_cases_match_cpu_dual_numbers(case, oracle_threads):
    # the gradient of a random weighting of every wall's reaction history,
    # against the oracle on dual numbers (the same frozen-branch map)
    import oracle
    from perigrain_amd import Stepper
    s, p, steps, ck, names, rng = _random_adjoint_case(case)
    print(f"case {case}: {s.dim}D, {s.n_nodes} nodes, {steps} steps, checkpoints every {ck}, "
          f"{'+'.join(names) or 'plain'}", flush=True)
    R, dR = oracle.reaction_history_tangent(s, p, steps, threads=oracle_threads)
    top = 2 if s.dim == 2 else 5
    with Stepper(s, p, device=0) as g:
        hist = g.adjoint_forward(steps, ck)
        if s.dim == 3:
            assert np.array_equal(hist, R), "the recorded 3D reaction history is bit-exact"
        else:
            np.testing.assert_allclose(hist, R, rtol=1e-10, atol=1e-10 * np.abs(R).max())
        assert np.any(hist[:, top] != 0), "the top wall should push"
        x0, v0, b0 = g.read("CurrPos"), g.read("vel"), g.bond_intact()
        if p.enable_fracture:
            assert np.any(b0 == 0), "the expanding grain should break bonds"
        bar = rng.normal(size=hist.shape) * np.abs(hist).max()
        raw = g.adjoint_backward(bar)
        assert np.array_equal(g.read("CurrPos"), x0) and np.array_equal(g.read("vel"), v0)
        assert np.array_equal(g.bond_intact(), b0)
    raw_cpu = np.einsum("tsk,tskq->q", bar, dR)
    # each parameter's tolerance scales with the sum of its terms' magnitudes,
    # so a sum that cancels, or a parameter the run does not depend on (μ
    # with friction off: zero on both sides), is still held to 1e-8
    mag = np.einsum("tsk,tskq->q", np.abs(bar), np.abs(dR))
    print("adjoint", raw, "cpu dual numbers", raw_cpu, "term magnitudes", mag,
          f"{int((b0 == 0).sum())} broken slots", flush=True)
    assert np.all(np.abs(raw - raw_cpu) <= 1e-8 * mag), (raw, raw_cpu, mag)


@pytest.mark.parametrize("pre", [0, 7])
def test_stored_trajectory_equals_checkpointed_recompute(pre):
    # One segment (checkpoint_every >= nsteps) keeps every step's state from
    # the recorded pass; several segments recompute them from checkpoints.
    # Both give the same reaction history and gradient (the latter to the
    # order in which the contact terms of different candidate lists are
    # summed), and the handle reads and steps on like a plain run.  pre:
    # plain steps before the recorded pass.
    from perigrain_amd import Stepper
    s, p = _fracturing_compression()
    steps = 120
    fields = ("CurrPos", "vel", "acc", "force")
    with Stepper(s, p, device=0) as ref:
        ref.step(pre + steps)
        want = {f: ref.read(f) for f in fields}
        want_wall = ref.wall()
        ref.step(30)
        want_next = ref.read("CurrPos")
    runs = []
    for ck in (30, 120, 500):  # 4 segments; one segment; one, with spare
        with Stepper(s, p, device=0) as g:
            if pre:
                g.step(pre)
            hist = g.adjoint_forward(steps, ck)
            for f in fields:
                assert np.array_equal(g.read(f), want[f]), f"ck={ck}: {f} after the recorded pass"
            w, r = g.wall()
            assert np.array_equal(w, want_wall[0]) and np.array_equal(r, want_wall[1]), ck
            bar = np.zeros_like(hist)
            bar[:, 2] = 2.0 * hist[:, 2]
            raw = g.adjoint_backward(bar)
            assert np.array_equal(g.read("CurrPos"), want["CurrPos"]), f"ck={ck} after backward"
            assert np.array_equal(g.read("vel"), want["vel"]), f"ck={ck} after backward"
            g.step(30)
            assert np.array_equal(g.read("CurrPos"), want_next), f"ck={ck}: stepping on"
            runs.append((ck, hist, raw))
    print([(ck, raw) for ck, _, raw in runs], flush=True)
    for ck, hist, raw in runs[1:]:
        assert np.array_equal(hist, runs[0][1]), ck
        np.testing.assert_allclose(raw, runs[0][2], rtol=1e-12)
