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


def cpu_baseline(circ, psi, seconds=12.0, seed=7):
    """Reference-faithful numpy restatement, single thread, bounded sample."""
    from oracle.reference_sv import ReferenceSV
    rng = np.random.default_rng(seed)
    sim = ReferenceSV(circ, psi)
    n = len(circ.trainable_nodes)
    sim.run(rng.uniform(0, 2 * np.pi, n), output_form="sv", input_state=psi)  # warm-up
    count, t0 = 0, time.perf_counter()
    while True:
        sim.run(rng.uniform(0, 2 * np.pi, n), output_form="sv", input_state=psi)
        count += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    return count / el, count, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1 << 21)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding over angle vectors is described in DESIGN.md)"}))
        sys.exit(2)

    import torch
    import mentpy_amd as mp
    from mentpy_amd.operators.gates import haar_unitary
    from mentpy_amd.simulators import sv_forward

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    rng = np.random.default_rng(2024)
    circ = mp.templates.grid_cluster(4, 20)
    sim = mp.PatternSimulator(circ, backend="hip-sv")
    plan = sim.simulator.plan
    n_ang = len(circ.trainable_nodes)
    psi = haar_unitary(16, rng)[:, 0]
    B = args.batch
    angles = torch.rand((B, n_ang), generator=torch.Generator().manual_seed(11),
                        dtype=torch.float32).mul_(2 * np.pi).to(dev)
    states = torch.as_tensor(psi, dtype=torch.complex64, device=dev).reshape(1, -1)
    out = torch.empty((B, 1, 16), dtype=torch.complex64, device=dev)

    def fwd():
        sv_forward(plan, angles, states, ref_phase=True, renormalize=True, out=out)

    secs = time_launches(fwd, args.steps, args.warmup, stream)
    ms = secs / args.steps * 1e3
    sims_per_s = B * args.steps / secs

    # c64 vs c128 error on a sample
    from oracle.vectorized import sv_batch
    idx = np.random.default_rng(5).choice(B, 512, replace=False)
    a_s = angles[idx].double().cpu().numpy()
    ref = sv_batch(circ, a_s, psi[None])[:, 0]
    got = out[idx, 0].cpu().numpy()
    amp_err = float(np.abs(got - ref).max())
    fid = np.abs(np.sum(np.conj(ref) * got, -1)) ** 2
    fid_err = float(np.max(np.abs(1 - fid)))

    f_sim = flops_per_sim(plan.info["n_steps"], plan.info["window"])
    achieved = f_sim * B / (ms / 1e3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "kernel": "k_sv_forward<4,true>", "flops_per_sim": f_sim}

    line = {
        "metric": "pattern sims/s + adjoint angle-grads/s, 4x20 grid, c64 vs c128 error",
        "value": sims_per_s, "unit": "sims/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic (seeded uniform angles, Haar input)",
        "config": {"workload": "grid_cluster(4,20) hip-sv forward, B angle vectors x 1 Haar input",
                   "batch": B, "n_angles": n_ang, "window": plan.info["window"]},
        "max_amp_err_vs_c128": amp_err, "max_infidelity_vs_c128": fid_err,
        "roofline": roofline,
    }
    if not args.no_cpu:
        rate, count, el = cpu_baseline(circ, psi, args.cpu_seconds)
        line["cpu_baseline"] = {"value": rate, "unit": "sims/s", "cores": 1, "kind": "port",
                                "sample": f"{count} sims of grid_cluster(4,20) in {el:.1f} s, "
                                          "step-by-step numpy complex128 restatement, 1 thread"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
