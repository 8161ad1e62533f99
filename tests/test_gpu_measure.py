"""Step-by-step measure() on the GPU against the oracle's step-by-step
restatement of NumpySimulatorSV.measure / NumpySimulatorDM.measure
(np_simulator_sv.py:164-225, np_simulator_dm.py:151-216): the window state
after every measurement, in the reference's qubit order and global phase.
Bar: max amplitude (entry) difference 2e-6 (complex64, normalised)."""
import numpy as np
import pytest

from conftest import haar_state

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _pattern(rows, cols):
    import mentpy_amd as mp
    g = mp.templates.linear_cluster(cols) if rows == 1 else mp.templates.grid_cluster(rows, cols)
    if rows > 1:
        g[g.outputc[1]] = mp.Ment("X")
    return g


@pytest.mark.parametrize("rows,cols,window", [(1, 5, 1), (2, 4, 1), (3, 5, 1), (2, 5, 4), (2, 4, 2),
                                              (4, 4, 1)])
def test_sv_measure_steps_match_oracle(rows, cols, window, rng):
    _torch()
    import mentpy_amd as mp
    from oracle.reference_sv import ReferenceSV
    g = _pattern(rows, cols)
    psi = haar_state(rows, rng)
    kw = {} if window == 1 else {"window_size": window, "schedule": list(g.measurement_order)}
    sim = mp.PatternSimulator(g, input_state=psi, backend="hip-sv", **kw)
    ref = ReferenceSV(g, psi, window_size=sim.simulator.window_size,
                      schedule=list(sim.simulator.schedule))
    worst = 0.0
    for node in sim.simulator.schedule_measure:
        m = g[node]
        th = float(rng.uniform(0, 2 * np.pi)) if m.angle is None else m.angle
        got, out = sim.measure(th)
        want, _ = ref.measure(th)
        assert out == 0 and got.shape == want.shape
        worst = max(worst, float(np.abs(got - want).max()))
    print(f"{rows}x{cols} w={window}: max window err {worst:.2e}", flush=True)
    assert worst < 2e-6
    with pytest.raises(ValueError, match="No more measurements"):
        sim.measure(0.1)


def test_dm_measure_steps_match_oracle(rng):
    _torch()
    import mentpy_amd as mp
    from oracle.reference_dm import ReferenceDM
    g = _pattern(3, 4)
    psi = haar_state(3, rng)
    sim = mp.PatternSimulator(g, input_state=psi, backend="hip-dm")
    ref = ReferenceDM(g, psi)
    worst = 0.0
    for node in sim.simulator.schedule_measure:
        m = g[node]
        th = float(rng.uniform(0, 2 * np.pi)) if m.angle is None else m.angle
        got, _ = sim.measure(th)
        want, _ = ref.measure(th)
        worst = max(worst, float(np.abs(got - want).max()))
    print(f"dm 3x4: max window err {worst:.2e}", flush=True)
    assert worst < 2e-6


def test_measure_api_errors_and_run_equivalence(rng):
    _torch()
    import mentpy_amd as mp
    g = _pattern(2, 4)
    psi = haar_state(2, rng)
    th = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes))
    sim = mp.PatternSimulator(g, input_state=psi, backend="hip-sv")
    full = sim.run(th, output_form="sv")
    sim.reset()
    tr = list(g.trainable_nodes)
    for node in sim.simulator.schedule_measure:
        m = g[node]
        if m.angle is not None:
            with pytest.raises(ValueError, match="fixed angle"):
                sim.measure(m.angle + 0.5)
        state, _ = sim.measure(th[tr.index(node)] if node in tr else m.angle)
    assert np.abs(state - full).max() < 2e-6           # last window = the output
    sim.reset()
    sim.measure(th[0])
    with pytest.raises(ValueError, match="No more measurements"):
        sim.run(th)
    sim.reset()
    with pytest.raises(ValueError, match="trainable"):
        sim.measure(None)
