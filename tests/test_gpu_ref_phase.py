"""Reference global phase at angles next to odd multiples of pi, on the GPU.

mentpy's amplitude-sum partial trace (`mentpy/calculator/state_ops.py:68-73`)
gives each step the factor sgn(cos(theta/2)) e^{i theta/2}, whose sign flips
at theta = (2k+1) pi; float32(pi) lies just above pi, so its factor is -i.
These angles go through every kernel that applies the reference phase (the
periodic and table forward kernels, the LDS-resident kernel for large
windows, the batched persistent kernel with and without shared sine/cosine
work, and the step-by-step window kernel) and the outputs must equal the
oracle's (`oracle/vectorized.py`, the literal amplitude sum in complex128)
without any phase alignment.
"""
import numpy as np
import pytest

from conftest import haar_state

pytestmark = pytest.mark.gpu

AMP_TOL = 2e-5


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _special_angles():
    pi32 = np.float32(np.pi)
    vals = [pi32, np.nextafter(pi32, np.float32(4)), np.nextafter(pi32, np.float32(3)),
            np.float32(-np.pi), np.float32(3 * np.pi), np.float32(5 * np.pi),
            np.float32(9.424778), np.float32(28.274334), np.float32(40.840706),
            np.float32(47.12389)]
    return np.array(vals, np.float32)


def _angles(rng, B, n):
    """Random angles with about a third replaced by the special values; row 0
    is float32(pi) everywhere."""
    th = rng.uniform(0, 2 * np.pi, (B, n)).astype(np.float32)
    sp = _special_angles()
    mask = rng.random((B, n)) < 0.35
    th[mask] = sp[rng.integers(0, sp.size, int(mask.sum()))]
    th[0, :] = np.float32(np.pi)
    return th


def _oracle(g, th, st, **kw):
    from oracle.vectorized import sv_batch
    return sv_batch(g, th.astype(np.float64), st, **kw)


@pytest.mark.parametrize("rows,cols,table", [(4, 20, False), (4, 20, True), (3, 7, False),
                                             (1, 9, False), (5, 6, True)])
def test_forward_reference_phase_near_odd_pi(rows, cols, table, rng):
    _torch()
    import mentpy_amd as mp
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    th = _angles(rng, 48, len(g.trainable_nodes))
    st = np.stack([haar_state(len(g.input_nodes), rng) for _ in range(3)])
    from mentpy_amd.simulators.hip_simulator import _as_angles, _as_states, sv_forward
    core = sim.simulator
    a = _as_angles(th, core.plan.info["n_angles"], core.device)
    s = _as_states(st, core.plan.info["n_inputs"], core.device)
    out = sv_forward(core.plan, a, s, table_steps=table).cpu().numpy()
    ref = _oracle(g, th, st)
    err = float(np.abs(out - ref).max())
    print(f"{rows}x{cols} table={table}: max amp err {err:.2e} (no phase alignment)", flush=True)
    assert err < AMP_TOL


def test_lds_kernel_reference_phase_near_odd_pi(rng):
    """Eight live qubits: the workgroup-per-simulation LDS kernel."""
    _torch()
    import mentpy_amd as mp
    g = mp.templates.grid_cluster(8, 3)
    sim = mp.PatternSimulator(g, backend="hip-sv")
    assert sim.simulator.plan.info["n_slots"] == 8
    th = _angles(rng, 6, len(g.trainable_nodes))
    st = np.stack([haar_state(8, rng) for _ in range(2)])
    out = sim.run_batch(th, st).cpu().numpy()
    err = float(np.abs(out - _oracle(g, th, st)).max())
    print(f"8x3 LDS kernel: max amp err {err:.2e}", flush=True)
    assert err < AMP_TOL


def test_batched_reference_phase_near_odd_pi(rng):
    """A job with S a multiple of the sharing group (lanes share the sine and
    cosine work) and one without."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.batched import BatchedForward
    jobs, refs = [], []
    for (r, c, S) in [(4, 20, 4), (3, 10, 1), (2, 12, 2)]:
        g = mp.templates.grid_cluster(r, c)
        sim = mp.PatternSimulator(g, backend="hip-sv")
        th = _angles(rng, 40, len(g.trainable_nodes))
        st = np.stack([haar_state(r, rng) for _ in range(S)])
        jobs.append((sim, th, st))
        refs.append(_oracle(g, th, st))
    bf = BatchedForward(jobs, sims_per_item=64)
    outs = bf.launch()
    torch.cuda.synchronize()
    for o, ref in zip(outs, refs):
        err = float(np.abs(o.cpu().numpy() - ref).max())
        print(f"batched job: max amp err {err:.2e}", flush=True)
        assert err < AMP_TOL


def test_window_kernel_reference_phase_near_odd_pi(rng):
    """Step-by-step measure() (window kernel) at float32(pi) and neighbours."""
    _torch()
    import mentpy_amd as mp
    from oracle.reference_sv import ReferenceSV
    g = mp.templates.grid_cluster(3, 5)
    psi = haar_state(3, rng)
    sim = mp.PatternSimulator(g, input_state=psi, backend="hip-sv")
    ref = ReferenceSV(g, psi)
    sp = _special_angles()
    worst = 0.0
    for j, node in enumerate(sim.simulator.schedule_measure):
        th = float(sp[j % sp.size])
        got, _ = sim.measure(th)
        want, _ = ref.measure(th)
        worst = max(worst, float(np.abs(got - want).max()))
    print(f"3x5 window kernel: max window err {worst:.2e}", flush=True)
    assert worst < 2e-6
