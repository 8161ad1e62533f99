"""dev_mode scheduling on the GPU against the step-by-step oracles in dev_mode
(`np_simulator_sv.py:176-197, 254-276`, `np_simulator_dm.py:151-216,
240-262`; the oracles are pinned on the CPU in tests/test_dev_mode.py).

Patterns whose dev-mode order leaves schedule order compile to plans off the
slot rotation: the SV forward runs on the LDS kernel, the DM forward on the
general kernel, windows on the window kernels; each is checked here against
the oracle (amplitudes with the reference phase, 2e-5; density matrices by
trace distance, 1e-5; windows after every step, 2e-6).  Entry points that
need schedule order refuse such plans.
"""
import numpy as np
import pytest

from conftest import haar_state
from test_dev_mode import CASES, _rows

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _sim(case, backend, **kw):
    import mentpy_amd as mp
    r, c, sched, w, order = CASES[case]
    g = mp.templates.grid_cluster(r, c)
    sim = mp.PatternSimulator(g, backend=backend, schedule=sched, window_size=w, dev_mode=True,
                              wires=_rows(r, c), **kw)
    assert sim.simulator.plan.info["rotating"] == 0
    return g, sim


@pytest.mark.parametrize("case", range(0, 40, 3))
def test_sv_dev_mode_run_matches_oracle(case, rng):
    _torch()
    from oracle.reference_sv import ReferenceSV
    g, sim = _sim(case, "hip-sv")
    r, c, sched, w, order = CASES[case]
    B = 12
    th = rng.uniform(-7, 7, (B, len(g.trainable_nodes))).astype(np.float32)
    psi = haar_state(r, rng)
    got = sim.simulator.run_batch(th, psi[None]).cpu().numpy()[:, 0]
    worst = 0.0
    for b in range(B):
        want = ReferenceSV(g, psi, window_size=w, schedule=sched, dev_mode=True,
                           wires=_rows(r, c)).run(th[b].astype(np.float64), output_form="sv")
        worst = max(worst, float(np.abs(got[b] - want).max()))
    print(f"case {case} ({r}x{c}, w={w}, lag {sim.simulator.plan.info['lag']}): "
          f"max amp err {worst:.2e}", flush=True)
    assert worst < 2e-5
    sim.reset(input_state=psi)
    sv = sim.run(th[0], output_form="sv")
    assert np.abs(sv - got[0]).max() < 1e-6
    assert list(sim.outcomes) == order


@pytest.mark.parametrize("case", [0, 7, 21, 33])
def test_sv_dev_mode_measure_steps(case, rng):
    _torch()
    from oracle.reference_sv import ReferenceSV
    g, sim = _sim(case, "hip-sv")
    r, c, sched, w, order = CASES[case]
    psi = haar_state(r, rng)
    sim.reset(input_state=psi)
    ref = ReferenceSV(g, psi, window_size=w, schedule=sched, dev_mode=True, wires=_rows(r, c))
    worst = 0.0
    for k, node in enumerate(order):
        th = float(np.float32(rng.uniform(0, 2 * np.pi)))
        assert sim.simulator.current_simulated_nodes() == ref.current_simulated_nodes()
        got, _ = sim.measure(th)
        want, _ = ref.measure(th)
        assert got.shape == want.shape
        worst = max(worst, float(np.abs(got - want).max()))
    print(f"case {case}: max window err {worst:.2e}", flush=True)
    assert worst < 2e-6
    assert list(sim.outcomes) == order


@pytest.mark.parametrize("case,p", [(1, 0.0), (6, 0.02), (13, 0.01), (29, 0.0), (35, 0.02)])
def test_dm_dev_mode_run_and_windows(case, p, rng):
    _torch()
    from oracle.reference_dm import ReferenceDM
    g, sim = _sim(case, "hip-dm", p_depol=p, p_readout=p / 2)
    r, c, sched, w, order = CASES[case]
    kw = dict(window_size=w, schedule=sched, dev_mode=True, wires=_rows(r, c), p_depol=p,
              p_readout=p / 2)
    B = 8
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes))).astype(np.float32)
    psi = haar_state(r, rng)
    got = sim.simulator.run_batch(th, psi[None]).cpu().numpy()[:, 0]
    td = 0.0
    for b in range(B):
        want = ReferenceDM(g, psi, **kw).run(th[b].astype(np.float64))
        td = max(td, float(0.5 * np.abs(np.linalg.eigvalsh(got[b] - want)).sum()))
    sim.reset(input_state=psi)
    ref = ReferenceDM(g, psi, **kw)
    worst = 0.0
    for node in order:
        a = float(np.float32(rng.uniform(0, 2 * np.pi)))
        got_w, out = sim.measure(a)
        want_w, wout = ref.measure(a)
        assert out == wout
        worst = max(worst, float(np.abs(got_w - want_w).max()))
    print(f"case {case} p={p}: trace distance {td:.2e}, max window err {worst:.2e}", flush=True)
    assert td < 1e-5 and worst < 2e-6


def test_dev_mode_refusals(rng):
    """Adjoint gradients, sampled outcomes and batched launches need schedule
    order; dev-mode grids with rows as wires keep it and run everywhere."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.batched import BatchedForward
    from mentpy_amd.gradients import AdjointEngine
    g, sim = _sim(0, "hip-sv")
    th = torch.rand((4, len(g.trainable_nodes)), device="cuda")
    r = CASES[0][0]
    cols = torch.eye(1 << r, dtype=torch.complex64, device="cuda")
    with pytest.raises(NotImplementedError):
        AdjointEngine(sim.simulator.plan).forward(th, cols)
    with pytest.raises(NotImplementedError):
        BatchedForward([(sim, th, haar_state(r, rng)[None])])
    with pytest.raises(NotImplementedError):
        _sim(0, "hip-sv", force0=False)
    grid = mp.templates.grid_cluster(3, 6)
    dev = mp.PatternSimulator(grid, backend="hip-sv", dev_mode=True, wires=_rows(3, 6))
    plain = mp.PatternSimulator(grid, backend="hip-sv")
    a = rng.uniform(0, 2 * np.pi, (16, len(grid.trainable_nodes))).astype(np.float32)
    st = haar_state(3, rng)[None]
    assert torch.equal(dev.simulator.run_batch(a, st), plain.simulator.run_batch(a, st))
