"""Benchmark: node-steps/s of the HIP PeriDEM stepper on the 3D compression
workload (64 spheres, ~265k nodes, ~23.5M directed bonds, fracture on,
moving top wall, gravity), one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched] [--no-adjoint]
                    [--no-configs] [--tuning SKIN,BLOCK]
    python bench.py --copy-calibration MIB   # plain device copy, for PMC scaling

Prints one JSON line.  `value` is node-steps/s with the state resident in
HBM; the timed region is K steps enqueued through pg_step, bracketed by HIP
events on the stepper's own stream.  The Verlet skin and the block size come
from the project's autotuner (pg_autotune), run first on a separate handle
of the same workload; both are result-neutral.  `roofline` prices the fused step
kernel against HBM bandwidth: the algorithmic bytes of one step,
N*(8*(8d+1)+16) + 5*M (N nodes, M directed bond slots, d=3), over the
step kernel's average duration, which libpg measures with HIP events
around its runs of step launches (pg_kernel_time).  `traffic` is the HBM
bytes per step-kernel launch from rocprofv3 FETCH_SIZE + WRITE_SIZE passes,
scaled by the same counters on a plain copy of known size; it is read from
profiles/step_kernel_pmc.json when that file was measured on this workload.
`cpu_baseline` times the FP64 CPU restatement of the reference algorithm
(the parity oracle) on a few steps of the same workload.

`adjoint` measures the differentiable form on the 2D compression of 64
grains (~15k nodes, 2000 steps, checkpoints every 100): the plain forward
run, the recorded forward pass and the adjoint (backward) pass, each timed on
the host clock after a warm-up run, and the adjoint gradient of the top-wall
reaction misfit against central finite differences (relative step 1e-6).

`other_workloads` times the two 2D workloads of the measurement plan that
fit one GPU beside the headline one (the reference's two-disk collision and
the 400-grain bulk settle) on the GPU, and the oracle on a sample of each.

`batched` measures the batched form on 256 independent 2D gravity-settle
instances (~2.3M nodes, uneven sizes): one handle holding all of them
(graph-replayed step chunks, autotuned skin and block) against the plain
per-instance loop on the same GPU (one handle per instance, same settings,
each advanced the same number of steps in turn).  Both are timed on the host
clock over whole pg_step calls.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "node-steps/s at 256k-node 3D; batched/plain speedup; adjoint\u00f7forward time"
PMC_FILE = os.path.join(ROOT, "profiles", "step_kernel_pmc.json")


def algorithmic_bytes(n_nodes, n_slots, dim):
    return n_nodes * (8 * (8 * dim + 1) + 16) + 5 * n_slots


def cpu_baseline(setup, params, threads, timed_steps=3, one_thread_grains=8):
    """The oracle (the reference algorithm restated in C++, OpenMP over
    grains) on the host: `timed_steps` steps of the whole workload after one
    warm-up step with `threads` threads, and one step with one thread on the
    workload's first `one_thread_grains` grains (a bounded sample: the whole
    workload takes about a minute per step on one thread)."""
    import oracle
    o = oracle.Oracle(setup, params, threads=threads)
    o.step(1)  # warm-up
    t0 = time.perf_counter()
    o.step(timed_steps)
    dt = time.perf_counter() - t0
    o.close()
    from perigrain_amd import setups
    sub = setups.subset(setup, list(range(min(one_thread_grains, setup.n_grains))))
    o1 = oracle.Oracle(sub, params, threads=1)
    o1.step(1)
    t0 = time.perf_counter()
    o1.step(1)
    dt1 = time.perf_counter() - t0
    o1.close()
    return {
        "value": setup.n_nodes * timed_steps / dt,
        "unit": "node-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{timed_steps} steps after 1 warm-up step of the same 3D workload "
                  f"({setup.n_nodes} nodes), {dt:.1f} s on {threads} threads",
        "one_thread": {"value": sub.n_nodes / dt1, "unit": "node-steps/s", "cores": 1,
                       "sample": f"1 step after 1 warm-up step of the workload's first "
                                 f"{sub.n_grains} grains ({sub.n_nodes} nodes), {dt1:.1f} s"},
    }


def copy_calibration(mib):
    """Copy a buffer of `mib` MiB device-to-device a few times (torch's copy
    kernel), so rocprofv3's FETCH_SIZE / WRITE_SIZE per dispatch can be
    compared with the known mib*1024 KiB read and written."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_kib": mib * 1024, "dispatches": 5}), flush=True)


def batched_bench(steps, seed, n_instances=256):
    from perigrain_amd import Stepper, setups
    t0 = time.perf_counter()
    S, P = setups.batched_settle(n_instances, seed)
    n_nodes = sum(s.n_nodes for s in S)
    print(f"bench: batched workload built in {time.perf_counter() - t0:.1f} s: "
          f"{n_instances} instances, {n_nodes} nodes", file=sys.stderr, flush=True)
    with Stepper(S, P, device=0) as b:
        b.set_graph(True)
        skin, block = b.autotune(trial_steps=64)
        b.step(64)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        t_batched = time.perf_counter() - t0
    print(f"bench: batched {steps} steps in {t_batched * 1e3:.1f} ms (skin {skin}, block {block})",
          file=sys.stderr, flush=True)
    handles = [Stepper(s, p, device=0) for s, p in zip(S, P)]
    try:
        for h in handles:
            h.set_graph(True)
            h.set_tuning(skin, block)
            h.step(64)
            h.sync()
        t0 = time.perf_counter()
        for h in handles:
            h.step(steps)
        for h in handles:
            h.sync()
        t_plain = time.perf_counter() - t0
    finally:
        for h in handles:
            h.close()
    print(f"bench: plain per-instance loop {steps} steps in {t_plain * 1e3:.1f} ms",
          file=sys.stderr, flush=True)
    return {
        "workload": f"{n_instances} x 2D gravity settle (uneven instances)",
        "nodes": n_nodes,
        "instances": n_instances,
        "steps": steps,
        "batched_node_steps_per_s": n_nodes * steps / t_batched,
        "plain_loop_node_steps_per_s": n_nodes * steps / t_plain,
        "speedup": t_plain / t_batched,
        "tuning": {"skin_factor": skin, "block": block, "graph": True},
    }


def adjoint_bench(steps=2000, ck_segments=100):
    from perigrain_amd import Stepper, diffrun
    run = diffrun.compression_2d_run(steps=steps)
    theta0, theta_t = (1.5e9, 1.0, 0.25), (1.0e9, 1.2, 0.4)
    target, st = run.forward(*theta_t)
    st[0].close()
    s, p = run.make_setup(*theta0)
    # every step's state kept (about 1.4 GB here), no recompute
    ck = diffrun.auto_checkpoint_every(s, steps)
    times = {}
    for rep in range(2):  # the first repetition warms up
        with Stepper(s, p, device=0) as g:
            g.sync()
            t0 = time.perf_counter()
            g.step(steps)
            g.sync()
            t_fwd = time.perf_counter() - t0
        with Stepper(s, p, device=0) as g:
            g.sync()
            t0 = time.perf_counter()
            hist = g.adjoint_forward(steps, ck)
            t_rec = time.perf_counter() - t0
            loss, bar = diffrun.top_wall_misfit(hist, target)
            t0 = time.perf_counter()
            raw = g.adjoint_backward(bar)
            t_bwd = time.perf_counter() - t0
        # the same with checkpoints every ck_segments steps, each segment
        # recomputed in the backward pass
        with Stepper(s, p, device=0) as g:
            g.sync()
            t0 = time.perf_counter()
            hist_c = g.adjoint_forward(steps, ck_segments)
            t_rec_c = time.perf_counter() - t0
            t0 = time.perf_counter()
            raw_seg = g.adjoint_backward(diffrun.top_wall_misfit(hist_c, target)[1])
            t_bwd_c = time.perf_counter() - t0
        times = {"forward_ms": t_fwd * 1e3, "recorded_forward_ms": t_rec * 1e3,
                 "backward_ms": t_bwd * 1e3}
        checkpointed = {"checkpoint_every": ck_segments, "recorded_forward_ms": t_rec_c * 1e3,
                        "backward_ms": t_bwd_c * 1e3,
                        "gradient_rel_diff": [abs(a - b) / abs(b) if b else abs(a)
                                              for a, b in zip(raw_seg, raw)]}
    Kn = s.normal_stiffness
    grad = [(raw[0] + Kn * raw[1]) / theta0[0], Kn * raw[1] / theta0[1], raw[2]]
    fd = []
    for k in range(3):
        h = 1e-6 * theta0[k]
        vals = []
        for sgn in (1, -1):
            th = list(theta0)
            th[k] += sgn * h
            hh, stt = run.forward(*th)
            stt[0].close()
            vals.append(diffrun.top_wall_misfit(hh, target)[0])
        fd.append((vals[0] - vals[1]) / (2 * h))
    rel = [abs(a - b) / abs(b) for a, b in zip(grad, fd)]
    # FP64 CPU reference gradient: the oracle on dual numbers (same frozen
    # branches), contracted with its own misfit cotangent
    import numpy as np
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    Rc, dRc = oracle.reaction_history_tangent(s, p, steps, threads=threads)
    t_cpu = time.perf_counter() - t0
    bar_c = np.zeros_like(Rc)
    bar_c[:, 2] = 2.0 * (Rc[:, 2] - target[:, 2])
    raw_c = np.einsum("tsk,tskq->q", bar_c, dRc)
    grad_c = [(raw_c[0] + Kn * raw_c[1]) / theta0[0], Kn * raw_c[1] / theta0[1], raw_c[2]]
    rel_c = [abs(a - b) / abs(b) for a, b in zip(grad, grad_c)]
    print(f"bench: adjoint forward {times['forward_ms']:.1f} ms, recorded "
          f"{times['recorded_forward_ms']:.1f} ms, backward {times['backward_ms']:.1f} ms "
          f"(checkpoints every {ck_segments}: recorded {checkpointed['recorded_forward_ms']:.1f} ms, "
          f"backward {checkpointed['backward_ms']:.1f} ms), "
          f"FD rel err {rel}, CPU dual-number rel err {rel_c} ({t_cpu:.1f} s)",
          file=sys.stderr, flush=True)
    return {
        "workload": "2D compression, 64 grains, top wall -0.5 m/s",
        "nodes": s.n_nodes,
        "steps": steps,
        "checkpoint_every": ck,
        **times,
        "checkpointed": checkpointed,
        "adjoint_over_forward": times["backward_ms"] / times["forward_ms"],
        "recorded_plus_adjoint_over_forward":
            (times["recorded_forward_ms"] + times["backward_ms"]) / times["forward_ms"],
        "theta": {"K": theta0[0], "contact_scale": theta0[1], "mu": theta0[2]},
        "loss": loss,
        "gradient": grad,
        "central_difference": fd,
        "fd_rel_err": rel,
        "cpu_dual_gradient": grad_c,
        "cpu_dual_rel_err": rel_c,
        "cpu_dual_threads": threads,
    }


def secondary_bench(threads, with_cpu):
    """The other single-GPU workloads of the measurement plan: the
    reference's own two-disk collision (2D, ~1.4k nodes, its full 5000
    steps) and the 2D bulk settle of 400 grains (~122k nodes, 2000 steps),
    each timed on the host clock over whole pg_step calls, and the oracle
    on a bounded sample of each with `threads` threads and with one."""
    from perigrain_amd import Stepper, setups
    out = {}
    cases = (
        ("two_disk_collision_2d", lambda: setups.two_disk_collision(seed=0), 5000, 10, 1000),
        ("bulk_settle_2d_400_grains", lambda: setups.bulk_settle_2d(seed=0), 2000, 1, 20),
    )
    for name, make, gpu_steps, cpu_warm, cpu_steps in cases:
        s, p = make()
        with Stepper(s, p, device=0) as g:  # warm-up handle
            g.step(10)
            g.sync()
        with Stepper(s, p, device=0) as g:
            g.sync()
            t0 = time.perf_counter()
            g.step(gpu_steps)
            g.sync()
            dt = time.perf_counter() - t0
        rec = {"nodes": s.n_nodes, "directed_bonds": s.n_slots(), "gpu_steps": gpu_steps,
               "gpu_ms_per_step": dt * 1e3 / gpu_steps,
               "gpu_node_steps_per_s": s.n_nodes * gpu_steps / dt}
        if with_cpu:
            import oracle
            for th in sorted({threads, 1}, reverse=True):
                o = oracle.Oracle(s, p, threads=th)
                o.step(cpu_warm)
                t0 = time.perf_counter()
                o.step(cpu_steps)
                dtc = time.perf_counter() - t0
                o.close()
                rec[f"cpu_node_steps_per_s_{th}_threads"] = s.n_nodes * cpu_steps / dtc
            rec["cpu_sample"] = f"oracle, {cpu_steps} steps after {cpu_warm} warm-up steps"
        print(f"bench: {name}: {rec}", file=sys.stderr, flush=True)
        out[name] = rec
    return out


def dropin_bench(setup, params, steps=10000, modulo=1000):
    """The drop-in executable (bin/simulate3d) on the headline workload with
    a snapshot every `modulo` steps, against pure stepping of the same steps
    through pg_step: the snapshots go out from a writer thread while the
    steps go on, so the time loop should take about as long as pure
    stepping.  Runs in a temporary directory laid out like the reference's
    root, removed afterwards."""
    import shutil
    import subprocess
    import tempfile
    from perigrain_amd import Stepper, h5io
    p = dataclasses.replace(params, timesteps=steps, modulo=modulo)
    tmp = tempfile.mkdtemp(prefix="pg_dropin_")
    try:
        os.makedirs(os.path.join(tmp, "config"))
        os.makedirs(os.path.join(tmp, "data", "hdf5"))
        h5io.write_all_h5(os.path.join(tmp, "data", "hdf5", "all.h5"), setup)
        keys = h5io.wall_speed_keys(setup.dim, p.wall_speed[:2 * setup.dim])
        h5io.write_config(os.path.join(tmp, "config", "main.conf"), p, **keys)
        with Stepper(setup, p, device=0) as g:
            g.step(10)
            g.sync()
        with Stepper(setup, p, device=0) as g:
            g.sync()
            t0 = time.perf_counter()
            g.step(steps)
            g.sync()
            pure = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = subprocess.run([h5io.simulate_exe(setup.dim)], cwd=tmp, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("simulate3d failed: " + r.stdout[-2000:])
        loop = [float(x) for x in h5io.read_dataset(
            os.path.join(tmp, "output", "hdf5", "run_time.h5"), "run_time").ravel()]
        out = os.path.join(tmp, "output", "hdf5")
        written = sum(os.path.getsize(os.path.join(out, f)) for f in os.listdir(out))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    rec = {"steps": steps, "modulo": modulo, "snapshots": steps // modulo,
           "pure_stepping_s": pure, "time_loop_s": sum(loop) / 1e3,
           "time_loop_over_pure": sum(loop) / 1e3 / pure, "process_wall_s": wall,
           "bytes_written": written}
    print(f"bench: drop-in {rec}", file=sys.stderr, flush=True)
    return rec


def measured_traffic(workload, n_nodes, n_slots):
    """HBM bytes per step-kernel launch from the committed PMC summary, if
    it was measured on this exact workload; else None."""
    try:
        with open(PMC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if (rec.get("workload"), rec.get("nodes"), rec.get("directed_bonds")) != (
            workload, n_nodes, n_slots):
        return None
    return rec.get("traffic_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-form leg")
    ap.add_argument("--batched-steps", type=int, default=512)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-form leg")
    ap.add_argument("--tuning", default="", metavar="SKIN,BLOCK",
                    help="Verlet skin factor and block size instead of the autotuner's choice")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 2D two-disk and bulk-settle workloads")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in executable's snapshot-overlap leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    if args.gpus != 1:
        print(f"bench.py: the stepper drives one GPU; --gpus {args.gpus} is not supported.",
              file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: multi-rank launch is not supported; run with one process.",
              file=sys.stderr)
        sys.exit(2)

    from perigrain_amd import Stepper, setups

    t0 = time.perf_counter()
    setup, params = setups.compression_3d(seed=args.seed)
    print(f"bench: workload built in {time.perf_counter() - t0:.1f} s: {setup.n_nodes} nodes, "
          f"{setup.n_slots()} directed bonds", file=sys.stderr, flush=True)
    dim = setup.dim
    with Stepper(setup, params, device=0) as g:
        # Verlet skin and block size from the autotuner, run on a separate
        # handle of the same workload so that the timed steps are the same
        # ones.  The timed handle is set up first and the autotuner's handle
        # is closed only after the timed steps: the trial steps then keep
        # the GPU busy up to the warm-up, so the timed steps do not start at
        # the clocks of a GPU that idled through a host-side set-up or
        # teardown (the step kernel runs 5-10% slower for the first tens of
        # milliseconds of load after the GPU idled for a few milliseconds).
        tune = None
        if args.tuning:
            skin, block = float(args.tuning.split(",")[0]), int(args.tuning.split(",")[1])
        else:
            t0 = time.perf_counter()
            tune = Stepper(setup, params, device=0)
            skin, block = tune.autotune(trial_steps=128)
            print(f"bench: autotuned skin factor {skin}, block {block} in "
                  f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        try:
            g.set_tuning(skin, block)
            g.step(args.warmup)
            g.sync()
            st0 = g.stats()
            k0 = g.kernel_time()
            g.event(0)
            g.step(args.steps)
            g.event(1)
            ms = g.elapsed_ms(0, 1)
            g.sync()
            st1 = g.stats()
            k1 = g.kernel_time()
        finally:
            if tune is not None:
                tune.close()
    n_nodes, n_slots = setup.n_nodes, setup.n_slots()
    ms_per_step = ms / args.steps
    value = n_nodes * args.steps / (ms / 1e3)
    bytes_step = algorithmic_bytes(n_nodes, n_slots, dim)
    k_launches = k1[1] - k0[1]
    kernel_us = (k1[0] - k0[0]) * 1e3 / k_launches if k_launches else ms_per_step * 1e3
    achieved = bytes_step / (kernel_us / 1e6) / 1e9
    print(f"bench: {args.steps} steps in {ms:.2f} ms ({ms_per_step * 1e3:.2f} us/step), "
          f"step kernel {kernel_us:.2f} us over {k_launches} launches, "
          f"list rebuilds {st1['rebuilds'] - st0['rebuilds']}, "
          f"skipped steps {st1['skipped_steps'] - st0['skipped_steps']}",
          file=sys.stderr, flush=True)
    workload = "3d_compression_64_spheres"

    batched = None
    if not args.no_batched:
        batched = batched_bench(args.batched_steps, args.seed)

    adj = None
    if not args.no_adjoint:
        adj = adjoint_bench()

    dropin = None
    if not args.no_dropin:
        dropin = dropin_bench(setup, params)

    threads = args.cpu_threads or max(1, min(16, os.cpu_count() or 1))
    others = None
    if not args.no_configs:
        others = secondary_bench(threads, not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(setup, params, threads)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered-lattice spheres)",
        "config": {
            "workload": workload,
            "nodes": n_nodes,
            "directed_bonds": n_slots,
            "grains": setup.n_grains,
            "dt": params.dt,
            "fracture": bool(params.enable_fracture),
            "top_wall_speed": float(params.wall_speed[5]),
            "tuning": {"skin_factor": skin, "block": block},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(workload, n_nodes, n_slots),
            "algorithmic_bytes_per_launch": bytes_step,
            "kernel_us": kernel_us,
            "kernel_launches": k_launches,
        },
        "cpu_baseline": cpu,
        "batched": batched,
        "adjoint": adj,
        "other_workloads": others,
        "dropin": dropin,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
