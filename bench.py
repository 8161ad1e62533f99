"""Benchmark: pattern sims/s and adjoint angle-gradients/s on the 4x20 grid cluster.

``python bench.py [--gpus 1] [--steps K] [--warmup W]`` prints ONE JSON line.

Workload (synthetic, seeded): ``grid_cluster(4, 20)`` (80 qubits, 76 measured
angles, window 5), a batch of B angle vectors drawn uniform in [0, 2 pi) and
one Haar-random 4-qubit input; one "step" is one ``hip-sv`` forward launch
over the whole batch, inputs already resident in HBM.  ``value`` is
simulations per second.  The same line reports the adjoint angle-gradient
rate on the same grid (MB-QML infidelity loss over 256 Haar inputs, 1024
parameter sets) against the plain GPU finite-difference loop, the
complex64-vs-complex128 errors of a sample, the roofline of the dominant
kernel, and the CPU baseline (the step-by-step numpy restatement of the
reference, timed on a bounded sample on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3     # MI355X published FP32 (vector and matrix) peak
HBM_PEAK_GBS = 8000.0
# HBM traffic is measured with rocprofv3 counters (FETCH_SIZE, WRITE_SIZE per
# dispatch, scaled by a device copy of known size) and recorded in the profile
# summaries, each naming the kernel and workload it measured.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "sv_forward_grid4x20.md")
BATCHED_PROFILE = os.path.join(ROOT, "profiles", "batched_64grids.md")
DM_PROFILE = os.path.join(ROOT, "profiles", "dm_forward_grid3x12.md")
MBQML_PROFILE = os.path.join(ROOT, "profiles", "mbqml_step_grid4x16.md")


def _rel(path):
    return os.path.relpath(path, ROOT)


def profiled_traffic(kernel, profile=TRAFFIC_PROFILE, per="sim"):
    """HBM bytes per simulation (``per="sim"``) or per launch (``"launch"``)
    of ``kernel`` from a profile summary's ``traffic_bytes_per_<per> <kernel>
    <bytes>`` line (None if the summary is absent or measured another kernel)."""
    tag = f"traffic_bytes_per_{per}"
    try:
        with open(profile) as f:
            for line in f:
                parts = line.strip().strip("`").split()
                if len(parts) == 3 and parts[0] == tag and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def flops_per_sim(n_steps, window):
    """Algorithmic flops of one simulation: 14 * 2^(w-1) per step (complex64)."""
    return n_steps * 14 * 2 ** (window - 1)


def time_launches(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1e3


def time_each_launch(fn, steps, warmup, stream):
    """Mean duration of a single launch: an event pair around each of
    ``steps`` launches on ``stream`` (the stream the kernel runs on)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in ev])) / 1e3


def cpu_baseline(circ, psi, seconds=12.0, seed=7):
    """Reference-faithful numpy restatement, single thread, bounded sample."""
    from threadpoolctl import threadpool_limits
    from oracle.reference_sv import ReferenceSV
    rng = np.random.default_rng(seed)
    sim = ReferenceSV(circ, psi)
    n = len(circ.trainable_nodes)
    with threadpool_limits(limits=1):
        sim.run(rng.uniform(0, 2 * np.pi, n), output_form="sv", input_state=psi)  # warm-up
        count, t0 = 0, time.perf_counter()
        while True:
            sim.run(rng.uniform(0, 2 * np.pi, n), output_form="sv", input_state=psi)
            count += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                break
    return count / el, count, el


def _pool_worker(args):
    """One process of the multi-core CPU baseline: the same restatement on
    grid_cluster(4,20) for ``seconds``; returns the simulations it finished."""
    seconds, seed = args
    import mentpy_amd as mp
    from oracle.reference_sv import ReferenceSV
    rng = np.random.default_rng(seed)
    circ = mp.templates.grid_cluster(4, 20)
    psi = mp.utils.generate_haar_random_states(4, rng=rng)
    sim = ReferenceSV(circ, psi)
    n = len(circ.trainable_nodes)
    sim.run(rng.uniform(0, 2 * np.pi, n), output_form="sv", input_state=psi)
    count, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        sim.run(rng.uniform(0, 2 * np.pi, n), output_form="sv", input_state=psi)
        count += 1
    return count, time.perf_counter() - t0


# CPU share of one GPU on the benchmark hosts: worker pools are capped at 16
# processes whatever the visible core count is
POOL_CAP = 16


def cpu_pool_baseline(seconds=8.0, max_workers=POOL_CAP):
    """The restatement mapped over a process pool of the host's cores, one
    numpy thread each (as mentpy's parallel tutorial maps simulations over a
    pool, intro-to-mbqml-parallel.rst:43-47).  Workers are spawned fresh, so no
    GPU state is inherited."""
    import multiprocessing as mpr
    workers = max(1, min(max_workers, len(os.sched_getaffinity(0))))
    ctx = mpr.get_context("spawn")
    names = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
    saved = {k: os.environ.get(k) for k in names}
    os.environ.update({k: "1" for k in names})         # inherited by the spawned workers
    try:
        with ctx.Pool(workers) as pool:
            res = pool.map(_pool_worker, [(seconds, 100 + k) for k in range(workers)])
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return sum(c / t for c, t in res), workers, sum(c for c, _ in res)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def adjoint_section(circ, sim, dev, steps, stream, sets=(1024, 16384), S=256, seed=99):
    """Adjoint angle-gradients/s on the MB-QML infidelity loss (one graph-replayed
    step: adjoint forward, Gram-matrix loss/seed GEMM, adjoint backward) against
    the plain GPU finite-difference loop (2 batched forward launches per angle);
    gradient error vs the complex128 oracle's exact shift rule."""
    import torch
    from mentpy_amd.gradients import MBQMLStep, fd_loop_infidelity_grad
    from mentpy_amd.operators.gates import haar_unitary
    rng = np.random.default_rng(seed)
    n_in = len(circ.input_nodes)
    U = haar_unitary(2 ** n_in, rng)
    X = np.stack([haar_unitary(2 ** n_in, rng)[:, 0] for _ in range(S)])
    Y = X @ U.T
    n = len(circ.trainable_nodes)
    Xd = torch.as_tensor(X, dtype=torch.complex64, device=dev)
    Yd = torch.as_tensor(Y, dtype=torch.complex64, device=dev)
    res = {}
    for B in sets:
        th = torch.tensor(rng.uniform(0, 2 * np.pi, (B, n)), dtype=torch.float32, device=dev)
        row = {}
        for mode in ("gram", "direct"):
            step = MBQMLStep(sim, Xd, Yd, n_sets=B, mode=mode)
            step.angles.copy_(th)
            graph = step.capture()
            secs = time_launches(graph.replay, steps, 2, stream)
            row[mode] = B * steps / secs
            del graph, step
        for mode in ("direct", "gram"):
            secs = time_launches(lambda: fd_loop_infidelity_grad(sim, th, Xd, Yd, h=1e-3, mode=mode),
                                 1, 1, stream)
            row["fd_" + mode] = B / secs
        res[B] = row
    from oracle.gradients import exact_shift_gradient
    from oracle.vectorized import sv_batch
    th0 = torch.tensor(np.random.default_rng(seed + 1).uniform(0, 2 * np.pi, (sets[0], n)),
                       dtype=torch.float32, device=dev)
    step = MBQMLStep(sim, Xd, Yd, n_sets=sets[0], mode="gram")
    _, g = step.run(th0)
    t0 = th0[0].double().cpu().numpy()

    def cost(t):
        o = sv_batch(circ, t[None], X)[0]
        return 1.0 - np.mean(np.abs(np.sum(np.conj(Y) * o, -1)) ** 2)
    ref = exact_shift_gradient(cost, t0)
    err = float(np.abs(g[0].double().cpu().numpy() - ref).max())
    big, small = sets[-1], sets[0]
    out = {"adjoint_grads_per_s": res[big]["gram"], "fd_loop_grads_per_s": res[big]["fd_direct"],
           "adjoint_vs_fd_loop": res[big]["gram"] / res[big]["fd_direct"],
           "adjoint_direct_grads_per_s": res[big]["direct"],
           "fd_gram_grads_per_s": res[big]["fd_gram"],
           "adjoint_gram_vs_fd_gram": res[big]["gram"] / res[big]["fd_gram"],
           "adjoint_direct_vs_fd_direct": res[big]["direct"] / res[big]["fd_direct"],
           "adjoint_grads_per_s_1024": res[small]["gram"],
           "fd_loop_grads_per_s_1024": res[small]["fd_direct"],
           "adjoint_vs_fd_loop_1024": res[small]["gram"] / res[small]["fd_direct"],
           "adjoint_gram_vs_fd_gram_1024": res[small]["gram"] / res[small]["fd_gram"],
           "adjoint_direct_vs_fd_direct_1024": res[small]["direct"] / res[small]["fd_direct"],
           "grad_max_abs_err_vs_c128": err,
           "grad_workload": f"4x20 grid ({n} angles), loss 1-mean_s|<U psi_s|out_s>|^2 over {S} Haar "
                            f"inputs; adjoint = one graph-replayed step (transfer matrix + data Gram "
                            f"matrix GEMM, or the inputs propagated: direct); FD loops = 2 forward "
                            f"launches + 1 fused loss kernel per angle, propagating the inputs "
                            f"(fd_loop, the plain loop; direct) or the basis columns with the Gram "
                            f"loss (fd_gram); {big} parameter sets (and {small}); a grad is the full "
                            f"{n}-angle gradient of one parameter set"}
    return out, (X, Y, t0, cost)


def cpu_gradient_baseline(circ, X, Y, t0, seconds_per_eval_cap=30.0):
    """mentpy's central FD gradient with the step-by-step numpy restatement:
    2 cost evaluations (S sims each) of one angle timed, scaled to all angles."""
    from oracle.reference_sv import ReferenceSV
    sim = ReferenceSV(circ, X[0])
    n = len(circ.trainable_nodes)

    def cost(t):
        acc = 0.0
        for x, y in zip(X, Y):
            o = sim.run(t, output_form="sv", input_state=x)
            acc += 1 - abs(np.vdot(y, o)) ** 2
        return acc / len(X)
    e = np.zeros(n)
    e[0] = 1e-5
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1):
        tic = time.perf_counter()
        (cost(t0 + e) - cost(t0 - e)) / 2e-5
        el = time.perf_counter() - tic
    return 1.0 / (el * n), el


def config_batched(args, dev, stream):
    """64 grids (rows 2..5, cols 10..40) x 1024 angle vectors x 16 Haar inputs
    = 1,048,576 simulations in one autotuned, graph-captured persistent launch."""
    import torch
    import mentpy_amd as mp
    from mentpy_amd.batched import BatchedForward, autotune
    from mentpy_amd.operators.gates import haar_unitary
    from oracle.vectorized import sv_batch
    rng = np.random.default_rng(3)
    jobs, pats, flops = [], [], 0
    for _ in range(64):
        r, c = int(rng.integers(2, 6)), int(rng.integers(10, 41))
        g = mp.templates.grid_cluster(r, c)
        sim = mp.PatternSimulator(g, backend="hip-sv")
        n = len(g.trainable_nodes)
        th = torch.rand((n, 1024), device=dev) * (2 * np.pi)
        st = np.stack([haar_unitary(2 ** r, rng)[:, 0] for _ in range(16)])
        jobs.append((sim, th.t(), st))
        pats.append((g, th, st))
        flops += 1024 * 16 * flops_per_sim(sim.simulator.plan.info["n_steps"], r + 1)
    bf = BatchedForward(jobs)
    cfg = autotune(bf)
    graph = bf.capture_graph()
    secs = time_launches(graph.replay, args.steps, args.warmup, stream)
    ms = secs / args.steps * 1e3
    total = bf.total_sims
    errs = []
    for (g, th, st), out in zip(pats[:16], bf.outputs[:16]):
        rows = np.random.default_rng(1).choice(1024, 8, replace=False)
        ref = sv_batch(g, th.t()[rows].double().cpu().numpy(), st)
        errs.append(float(np.abs(out[rows].cpu().numpy() - ref).max()))
    achieved = flops / (ms / 1e3) / 1e12
    return {"metric": "batched pattern sims/s, 64 uneven grids x 1024 angles x 16 inputs",
            "value": total / (ms / 1e3), "unit": "sims/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded)",
            "config": {"workload": "64 grid_cluster(r in 2..5, c in 10..40), 1024 angle vectors x 16 "
                                   "Haar inputs each, one persistent graph-replayed launch",
                       "sims": total, "autotune": cfg},
            "max_amp_err_vs_c128": max(errs),
            "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                         "traffic": profiled_traffic("k_batch", BATCHED_PROFILE, per="launch"),
                         "traffic_note": "HBM bytes per graph replay (the four window-size "
                                         "kernels), rocprof counters in " + _rel(BATCHED_PROFILE),
                         "kernel": "k_batch"}}


def config_dm(args, dev, stream):
    """Density-matrix 3x12 grid, depolarizing 0.01 + readout flip 0.005, 4096 angle vectors."""
    import torch
    import mentpy_amd as mp
    from mentpy_amd.operators.gates import haar_unitary
    from mentpy_amd.simulators import dm_forward
    from oracle.vectorized import dm_batch
    rng = np.random.default_rng(12)
    g = mp.templates.grid_cluster(3, 12)
    sim = mp.PatternSimulator(g, backend="hip-dm", p_depol=0.01, p_readout=0.005)
    n = len(g.trainable_nodes)
    B = args.dm_batch
    th = torch.rand((n, B), device=dev) * (2 * np.pi)
    psi = haar_unitary(8, rng)[:, 0]
    st = torch.as_tensor(psi, dtype=torch.complex64, device=dev).reshape(1, -1)
    out = torch.empty((B, 1, 8, 8), dtype=torch.complex64, device=dev)
    plan = sim.simulator.plan

    def run():
        dm_forward(plan, th.t(), st, p_depol=0.01, p_readout=0.005, out=out)
    # a 4096-simulation launch is shorter than one Python call of the wrapper:
    # the K launches are captured in one HIP graph and replayed (the eager
    # call loop is reported beside it)
    eager_ms = time_launches(run, args.steps, args.warmup, stream) / args.steps * 1e3
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph):
            for _ in range(args.steps):
                run()
    torch.cuda.synchronize()
    secs = time_launches(graph.replay, 1, 2, stream)
    ms = secs / args.steps * 1e3
    k = min(B, 4096)
    ref = dm_batch(g, th.t()[:k].double().cpu().numpy(), psi[None], p_depol=0.01, p_readout=0.005)[:, 0]
    td = 0.5 * np.abs(np.linalg.eigvalsh(out[:k, 0].cpu().numpy() - ref)).sum(-1)
    f_sim = plan.info["n_steps"] * 20 * 4 ** (plan.info["window"] - 1)
    achieved = f_sim * B / (ms / 1e3) / 1e12
    dm_kernel = "k_dm_forward<3,true>" if B >= 32768 else "k_dm_forward_lds<3,16,4,true>"
    bps = profiled_traffic(dm_kernel, DM_PROFILE)
    dm_traffic = None if bps is None else round(bps * B)
    return {"metric": "density-matrix pattern sims/s, 3x12 grid, noisy, trace distance vs c128",
            "value": B / (ms / 1e3), "unit": "sims/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded)",
            "config": {"workload": "grid_cluster(3,12) hip-dm, depolarizing 0.01, readout flip 0.005",
                       "batch": B},
            "max_trace_distance_vs_c128": float(td.max()),
            "timing": f"{args.steps} launches captured in one HIP graph, replayed",
            "eager_ms_per_launch": eager_ms,
            "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
                         "traffic": dm_traffic,
                         "kernel": dm_kernel, "flops_per_sim": f_sim}}


def config_sample(args, dev, stream):
    """Sampled outcomes with flow corrections (force0=False), 4x20 grid,
    2^20 angle vectors x 1 Haar input: every outcome drawn, Pauli frame tracked,
    output frame applied by a second pass."""
    import torch
    import mentpy_amd as mp
    from mentpy_amd.operators.gates import haar_unitary
    from mentpy_amd.simulators.hip_simulator import decode_outcomes, sv_forward, sv_sample
    rng = np.random.default_rng(21)
    g = mp.templates.grid_cluster(4, 20)
    sim = mp.PatternSimulator(g, backend="hip-sv", force0=False)
    plan = sim.simulator.plan
    n = len(g.trainable_nodes)
    B = args.sample_batch
    th = torch.rand((n, B), device=dev) * (2 * np.pi)
    st = torch.as_tensor(haar_unitary(16, rng)[:, 0], dtype=torch.complex64, device=dev).reshape(1, -1)
    out = torch.empty((B, 1, 16), dtype=torch.complex64, device=dev)
    W = (plan.info["n_steps"] + 31) // 32
    rows = torch.empty((B, 1, W + 2), dtype=torch.int32, device=dev)
    seed = [1]

    def run():
        seed[0] += 1
        sv_sample(plan, th.t(), st, seed[0], out=out, outcomes=rows)
    secs = time_launches(run, args.steps, args.warmup, stream)
    ms = secs / args.steps * 1e3
    post = sv_forward(plan, th.t()[:65536], st, ref_phase=False)
    torch.cuda.synchronize()
    ov = (post[:, 0].conj() * out[:65536, 0]).sum(-1).abs() ** 2
    bits = decode_outcomes(rows[:65536], plan.info["n_steps"])
    return {"metric": "sampled-outcome pattern sims/s (force0=False, flow corrections), 4x20 grid",
            "value": B / (ms / 1e3), "unit": "sims/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded)",
            "config": {"workload": "grid_cluster(4,20) hip-sv, force0=False, B angle vectors x 1 Haar input",
                       "batch": B, "kernels": "k_sv_sample<4,true> (periodic pair coefficients; output frame applied in-kernel)"},
            "max_infidelity_vs_postselected": float((1 - ov).abs().max().item()),
            "outcome_mean": float(bits.mean())}


def config_mbqml(args, dev, stream):
    """MB-QML training step: 4x16 grid (60 angles), 1024 parameter sets x 256 Haar inputs."""
    import torch
    import mentpy_amd as mp
    from mentpy_amd.gradients import MBQMLStep, fd_loop_infidelity_grad
    from mentpy_amd.operators.gates import haar_unitary
    from oracle.gradients import exact_shift_gradient
    from oracle.vectorized import sv_batch
    rng = np.random.default_rng(16)
    g = mp.templates.grid_cluster(4, 16)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    U = haar_unitary(16, rng)
    X = np.stack([haar_unitary(16, rng)[:, 0] for _ in range(256)])
    Y = X @ U.T
    B = 1024
    th = torch.tensor(rng.uniform(0, 2 * np.pi, (B, 60)), dtype=torch.float32, device=dev)
    step = MBQMLStep(sim, X, Y, n_sets=B, mode="gram")
    step.angles.copy_(th)
    graph = step.capture()
    secs = time_launches(graph.replay, args.steps, args.warmup, stream)
    ms = secs / args.steps * 1e3
    # roofline of the largest part of the step, the segmented forward (the
    # segment sweeps and the chain products), timed alone with HIP events
    fwd_s = time_each_launch(step._launch_forward, args.steps, args.warmup, stream)
    info = sim.simulator.plan.info
    n_cols = step.cols.shape[0]
    f_sweep = B * n_cols * flops_per_sim(info["n_steps"], info["window"])
    f_chain = B * (step.segments - 1) * n_cols ** 3 * 8
    fwd_flops = f_sweep + f_chain
    fwd_kernel = "k_sv_seg_forward<4,true>+k_seg_chain_forward_mfma"
    fwd_tflops = fwd_flops / fwd_s / 1e12
    fd_secs = time_launches(lambda: fd_loop_infidelity_grad(sim, th, step.X, step.Y, h=1e-3), 1, 1, stream)
    _, gr = step.run(th)
    # a full training iteration (adjoint step + device Adam) as one graph replay
    from mentpy_amd.optimizers import MBQMLTrainer
    trainer = MBQMLTrainer(sim, step.X, step.Y, th, step_size=0.05, mode="gram")
    trainer.capture()
    it_secs = time_launches(trainer._graph.replay, args.steps, args.warmup, stream)

    def cost(t):
        o = sv_batch(g, t[None], X)[0]
        return 1.0 - np.mean(np.abs(np.sum(np.conj(Y) * o, -1)) ** 2)
    ref = exact_shift_gradient(cost, th[0].double().cpu().numpy())
    err = float(np.abs(gr[0].double().cpu().numpy() - ref).max())
    return {"metric": "MB-QML adjoint angle-grads/s, 4x16 grid, 1024 sets x 256 inputs",
            "value": B / (ms / 1e3), "unit": "grads/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded Haar data, Haar target U)",
            "config": {"workload": "grid_cluster(4,16), 60 angles, loss 1-mean|<U psi|out>|^2",
                       "parameter_sets": B, "inputs": 256},
            "fd_loop_grads_per_s": B / fd_secs, "adjoint_vs_fd_loop": (B / (ms / 1e3)) / (B / fd_secs),
            "train_iterations_per_s": args.steps / it_secs,
            "train_note": "one graph replay = loss+gradient of all 1024 parameter sets + device Adam update",
            "grad_max_abs_err_vs_c128": err,
            "roofline": {"bound": "valu", "achieved": round(fwd_tflops, 3), "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(fwd_tflops / FP32_PEAK_TFLOPS, 4),
                         "traffic": profiled_traffic(fwd_kernel, MBQML_PROFILE, per="launch"),
                         "kernel": fwd_kernel, "kernel_ms": fwd_s * 1e3,
                         "flops": fwd_flops, "segments": step.segments}}


def config_copy(args, dev, stream):
    """HBM calibration probe: device-to-device copy of a 1 GiB buffer (the
    reference kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE counters)."""
    import torch
    n = (1 << 30) // 4
    x = torch.rand(n, device=dev)
    y = torch.empty_like(x)
    secs = time_launches(lambda: y.copy_(x), args.steps, args.warmup, stream)
    ms = secs / args.steps * 1e3
    return {"metric": "device copy GB/s (1 GiB read + 1 GiB write)", "value": 2 * 4 * n / (ms / 1e3) / 1e9,
            "unit": "GB/s", "ms_per_step": ms, "bytes_read": 4 * n, "bytes_written": 4 * n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1 << 21)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-pool-seconds", type=float, default=6.0)
    ap.add_argument("--grad-sets", type=int, nargs="+", default=[1024, 16384])
    ap.add_argument("--grad-inputs", type=int, default=256)
    ap.add_argument("--cpu-grad-inputs", type=int, default=32)
    ap.add_argument("--config", default="grid4x20",
                    choices=["grid4x20", "batched", "dm", "mbqml", "sample", "copy"])
    ap.add_argument("--sample-batch", type=int, default=1 << 20)
    ap.add_argument("--no-grad", action="store_true", help="skip the adjoint section (profiling runs)")
    ap.add_argument("--dm-batch", type=int, default=4096)
    ap.add_argument("--angles", default="per-sim", choices=["per-sim", "shared"],
                    help="headline workload variant for attribution profiles: one angle vector "
                         "per simulation (the benchmark) or one vector shared by all")
    ap.add_argument("--phase", default="reference", choices=["reference", "raw"],
                    help="output phase convention of the headline forward (reference: mentpy's)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 (simulations "
                          "are independent, so several GPUs would each take a share of the "
                          "angle vectors)"}))
        sys.exit(2)

    import torch
    import mentpy_amd as mp
    from mentpy_amd.operators.gates import haar_unitary
    from mentpy_amd.simulators import sv_forward

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    if args.config != "grid4x20":
        fn = {"batched": config_batched, "dm": config_dm, "mbqml": config_mbqml,
              "sample": config_sample, "copy": config_copy}[args.config]
        print(json.dumps(fn(args, dev, stream)), flush=True)
        return
    rng = np.random.default_rng(2024)
    circ = mp.templates.grid_cluster(4, 20)
    sim = mp.PatternSimulator(circ, backend="hip-sv")
    plan = sim.simulator.plan
    n_ang = len(circ.trainable_nodes)
    psi = haar_unitary(16, rng)[:, 0]
    B = args.batch
    # angle-major storage [n_angles, B] (each step's angles are contiguous across
    # the batch, so every angle load of a wavefront is one coalesced 256-byte read);
    # the kernels see it as the [B, n_angles] view angles_t.t()
    angles_t = torch.rand((n_ang, B), generator=torch.Generator().manual_seed(11),
                          dtype=torch.float32).mul_(2 * np.pi).to(dev)
    # the error sample always holds simulations at float32(pi), where the
    # reference phase sgn(cos(theta/2)) e^{i theta/2} is -i (float32(pi) > pi):
    # one with every angle at float32(pi), one with a few such angles
    idx = np.random.default_rng(5).choice(B, 512, replace=False)
    pi32 = float(np.float32(np.pi))
    angles_t[:, int(idx[0])] = pi32
    angles_t[::7, int(idx[1])] = pi32
    angles = angles_t.t()
    if args.angles == "shared":
        angles = angles_t[:, int(idx[2]):int(idx[2]) + 1].t().expand(B, n_ang)   # angle_stride 0
    ref_phase = args.phase == "reference"
    states = torch.as_tensor(psi, dtype=torch.complex64, device=dev).reshape(1, -1)
    out = torch.empty((B, 1, 16), dtype=torch.complex64, device=dev)

    def fwd():
        sv_forward(plan, angles, states, ref_phase=ref_phase, renormalize=True, out=out)

    secs = time_launches(fwd, args.steps, args.warmup, stream)
    ms = secs / args.steps * 1e3
    sims_per_s = B * args.steps / secs
    kernel_s = time_each_launch(fwd, args.steps, 1, stream)

    # single-simulation latency: one angle vector per launch (inputs in HBM),
    # and the host call run(angles) with its copies to and from the device
    one = angles[:1]
    out1 = torch.empty((1, 1, 16), dtype=torch.complex64, device=dev)
    lat_s = time_each_launch(lambda: sv_forward(plan, one, states, out=out1), 50, 5, stream)
    sim.reset(input_state=psi)
    th_host = one[0].double().cpu().numpy()
    sim.run(th_host, output_form="sv")
    t0 = time.perf_counter()
    for _ in range(20):
        sim.reset(input_state=psi)
        sim.run(th_host, output_form="sv")
    host_call_s = (time.perf_counter() - t0) / 20

    # c64 vs c128 error on a sample
    from oracle.vectorized import sv_batch
    a_s = angles[idx].double().cpu().numpy()
    ref = sv_batch(circ, a_s, psi[None], phase=args.phase)[:, 0]
    got = out[idx, 0].cpu().numpy()
    if ref_phase:
        amp_err = float(np.abs(got - ref).max())
    else:   # raw frame: the global phase is not fixed; compare after aligning it
        ov = np.sum(np.conj(ref) * got, -1)
        amp_err = float(np.abs(got - (ov / np.abs(ov))[:, None] * ref).max())
    fid = np.abs(np.sum(np.conj(ref) * got, -1)) ** 2
    fid_err = float(np.max(np.abs(1 - fid)))

    f_sim = flops_per_sim(plan.info["n_steps"], plan.info["window"])
    achieved = f_sim * B / kernel_s / 1e12
    r = "true" if ref_phase else "false"
    kernel = (f"k_sv_forward_periodic<4,{r}>" if plan.info["periodic"]
              else f"k_sv_forward<4,{r},true>")
    tps = profiled_traffic(kernel)
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": None if tps is None else int(tps * B),
                "kernel": kernel, "flops_per_sim": f_sim, "kernel_ms": kernel_s * 1e3,
                "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE per dispatch, "
                                  "profiles/sv_forward_grid4x20.md"}

    line = {
        "metric": "pattern sims/s + adjoint angle-grads/s, 4x20 grid, c64 vs c128 error",
        "value": sims_per_s, "unit": "sims/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded uniform angles, Haar input)",
        "config": {"workload": "grid_cluster(4,20) hip-sv forward, B angle vectors x 1 Haar input",
                   "batch": B, "n_angles": n_ang, "window": plan.info["window"],
                   "angle_layout": "angle-major [n_angles, B]"
                                   + ("" if args.angles == "per-sim" else
                                      " (attribution variant: one vector shared by all)"),
                   "phase": args.phase},
        "max_amp_err_vs_c128": amp_err, "max_infidelity_vs_c128": fid_err,
        "roofline": roofline,
        "single_sim_latency_us": lat_s * 1e6,
        "single_sim_host_call_us": host_call_s * 1e6,
        "latency_note": "one angle vector per launch with the input in HBM; host call = "
                        "reset + run(angles) with host arrays (copies in and out, sync)",
    }
    X = None
    if not args.no_grad:
        adj, (X, Y, t0, _) = adjoint_section(circ, sim, dev, max(2, args.steps // 4), stream,
                                             sets=tuple(args.grad_sets), S=args.grad_inputs)
        line.update(adj)
    if not args.no_cpu:
        rate, count, el = cpu_baseline(circ, psi, args.cpu_seconds)
        line["cpu_baseline"] = {"value": rate, "unit": "sims/s", "cores": 1, "kind": "port",
                                "sample": f"{count} sims of grid_cluster(4,20) in {el:.1f} s, "
                                          "step-by-step numpy complex128 restatement, 1 thread",
                                "cpu": cpu_model()}
        if args.cpu_pool_seconds > 0:
            prate, workers, pcount = cpu_pool_baseline(args.cpu_pool_seconds)
            line["cpu_baseline_pool"] = {
                "value": prate, "unit": "sims/s", "cores": workers, "kind": "port",
                "host_cores_visible": len(os.sched_getaffinity(0)), "pool_cap": POOL_CAP,
                "sample": f"{pcount} sims of grid_cluster(4,20), {workers} spawned processes x "
                          f"{args.cpu_pool_seconds:.0f} s, 1 numpy thread each (pool capped at "
                          f"{POOL_CAP}, the host's CPU share per GPU)"}
        if X is not None:
            grate, gel = cpu_gradient_baseline(circ, X[:args.cpu_grad_inputs], Y[:args.cpu_grad_inputs], t0)
            grate = grate * args.cpu_grad_inputs / len(X)
            line["cpu_baseline_grads"] = {
                "value": grate, "unit": "grads/s", "cores": 1, "kind": "port",
                "sample": f"central FD of one angle over {args.cpu_grad_inputs} of the {len(X)} "
                          f"inputs ({gel:.1f} s), scaled linearly to all angles and inputs"}
            line["adjoint_vs_cpu"] = line["adjoint_grads_per_s"] / grate
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
