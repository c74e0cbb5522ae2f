"""Batched-form machinery on a GPU: graph replay of step chunks and the
autotuner leave every result bit-identical to plain per-step launches."""
import numpy as np
import pytest

from perigrain_amd import Stepper, setups

pytestmark = pytest.mark.gpu

FIELDS = ("CurrPos", "vel", "acc", "force")


def _fast_settle():
    s, p = setups.bulk_settle_2d(seed=6, n_grains=16, box=(8e-3, 16e-3), nodes_per_grain=80)
    rng = np.random.default_rng(0)
    for gi in range(s.n_grains):
        a = rng.uniform(0, 2 * np.pi)
        s.vel[s.grain_nodes(gi)] = 100.0 * np.array([np.cos(a), np.sin(a)])
    return s, p


def _same(a, b, insts=(0,)):
    for k in insts:
        for f in FIELDS:
            assert np.array_equal(a.read(f, inst=k), b.read(f, inst=k)), (k, f)
        assert np.array_equal(a.bond_intact(inst=k), b.bond_intact(inst=k)), k
        assert np.array_equal(a.contacts(inst=k), b.contacts(inst=k)), k
        wa, ra = a.wall(inst=k)
        wb, rb = b.wall(inst=k)
        assert np.array_equal(wa, wb) and np.array_equal(ra, rb), k


def test_graph_replay_equals_launches_with_rebuilds():
    s, p = _fast_settle()
    with Stepper(s, p, device=0) as plain, Stepper(s, p, device=0) as graph:
        graph.set_graph(True)
        for n in (300, 1, 320):  # replayed chunks, a lone step, replayed chunks again
            plain.step(n)
            graph.step(n)
            _same(plain, graph)
        assert graph.stats()["rebuilds"] > 1, "the workload should force list rebuilds"


def test_graph_replay_batched_instances():
    sp = [setups.settle_instance_2d(seed) for seed in (21, 22)]
    ss = [setups.subset(s, list(range(min(10, s.n_grains)))) for s, _ in sp]
    ps = [p for _, p in sp]
    with Stepper(ss, ps, device=0) as plain, Stepper(ss, ps, device=0) as graph:
        graph.set_graph(True)
        plain.step(200)
        graph.step(200)
        _same(plain, graph, insts=(0, 1))


def test_autotune_is_result_neutral():
    s, p = _fast_settle()
    with Stepper(s, p, device=0) as tuned, Stepper(s, p, device=0) as plain:
        skin, block = tuned.autotune(trial_steps=20)
        assert skin in (0.25, 0.5, 1.0) and block in (64, 128, 256)
        assert tuned.t == 200  # a warm-up run and nine trials
        plain.step(200)
        _same(plain, tuned)
        tuned.step(150)
        plain.step(150)
        _same(plain, tuned)


def test_autotune_choice_is_persisted_per_workload(tmp_path, monkeypatch):
    # with PG_AUTOTUNE_CACHE set, a choice recorded for the same workload is
    # applied without trials; another workload gets its own entry
    cache = tmp_path / "autotune.txt"
    monkeypatch.setenv("PG_AUTOTUNE_CACHE", str(cache))
    s, p = _fast_settle()
    with Stepper(s, p, device=0) as a:
        choice = a.autotune(trial_steps=8)
        assert a.t == 80
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and "dim=2" in lines[0]
    # the recorded choice is reused by a new handle of the same workload: no steps run
    cache.write_text(lines[0].split(" | ")[0] + " | 1 64\n")
    with Stepper(s, p, device=0) as b:
        assert b.autotune(trial_steps=8) == (1.0, 64)
        assert b.t == 0
    s2, p2 = setups.two_disk_collision(seed=1)
    with Stepper(s2, p2, device=0) as c:
        c.autotune(trial_steps=4)
        assert c.t == 40
    assert len(cache.read_text().splitlines()) == 2
    assert choice[0] in (0.25, 0.5, 1.0)

