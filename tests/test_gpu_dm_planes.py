"""hip-dm in every measurement plane against the complex128 oracles (GPU).

mentpy's density-matrix backend measures any plane of ``Ment.matrix``
(`mentpy/simulators/np_simulator_dm.py:307-346`, `mentpy/operators/ment.py:218-253`):
XZ / YZ nodes have outcome probabilities other than 1/2, the outcome is 1
where outcome 0 has probability below 1e-4, and Z nodes in
``mode="expectation"`` are traced out unprojected.  The per-step-probability
kernel (``k_dm_general``) is checked against ``oracle/vectorized.py::dm_batch``
(pinned on the CPU against the step-by-step oracle and the brute force) on
mixed-plane grids with and without noise.  Bar: trace distance < 1e-5 and
max entry error < 2e-5 (complex64).
"""
import numpy as np
import pytest

from conftest import haar_state

pytestmark = pytest.mark.gpu

TD_TOL = 1e-5
ENTRY_TOL = 2e-5


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _trace_distance(a, b):
    return 0.5 * np.abs(np.linalg.eigvalsh(a - b)).sum(-1)


def _mixed(rows, cols, rng, z_nodes=0):
    import mentpy_amd as mp
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    meas = [v for v in g.measurement_order if v not in g.output_nodes]
    for v in meas:
        g[v] = mp.Ment(plane=str(rng.choice(["XY", "XZ", "YZ"])))
    picks = rng.choice(meas, 2 + z_nodes, replace=False)
    g[int(picks[0])] = mp.Ment("X")
    g[int(picks[1])] = mp.Ment(float(rng.uniform(0, 6)), "YZ")
    for v in picks[2:]:
        g[int(v)] = mp.Ment("Z")
    return g


@pytest.mark.parametrize("rows,cols,p", [(2, 3, 0.0), (3, 4, 0.0), (3, 4, 0.02), (3, 12, 0.0),
                                         (3, 12, 0.01), (1, 6, 0.0), (4, 5, 0.01)])
def test_dm_mixed_planes_match_oracle(rows, cols, p, rng):
    _torch()
    import mentpy_amd as mp
    from oracle.vectorized import dm_batch
    g = _mixed(rows, cols, rng)
    sim = mp.PatternSimulator(g, backend="hip-dm", p_depol=p, p_readout=p / 2)
    assert sim.simulator.plan.info["general_planes"] == 1
    B, S = 24, 2
    th = rng.uniform(0, 2 * np.pi, (B, len(g.trainable_nodes)))
    st = np.stack([haar_state(len(g.input_nodes), rng) for _ in range(S)])
    out = sim.run_batch(th, st).cpu().numpy()
    ref = dm_batch(g, th.astype(np.float32).astype(np.float64), st, p_depol=p, p_readout=p / 2)
    td = float(_trace_distance(out, ref).max())
    err = float(np.abs(out - ref).max())
    print(f"{rows}x{cols} p={p}: trace distance {td:.2e}, max entry err {err:.2e}", flush=True)
    assert td < TD_TOL and err < ENTRY_TOL


@pytest.mark.parametrize("rows,cols", [(2, 4), (3, 6)])
def test_dm_z_nodes_in_expectation_mode(rows, cols, rng):
    torch = _torch()
    import mentpy_amd as mp
    from oracle.reference_dm import ReferenceDM
    from oracle.vectorized import dm_batch
    g = _mixed(rows, cols, rng, z_nodes=1)
    sim = mp.PatternSimulator(g, backend="hip-dm", p_depol=0.01)
    th = rng.uniform(0, 2 * np.pi, (8, len(g.trainable_nodes)))
    st = haar_state(rows, rng)[None]
    with pytest.raises(NotImplementedError):
        sim.run_batch(th, st)
    out = sim.run_batch(th, st, mode="expectation").cpu().numpy()
    th32 = th.astype(np.float32).astype(np.float64)
    ref = dm_batch(g, th32, st, p_depol=0.01, mode="expectation")
    td = float(_trace_distance(out, ref).max())
    print(f"{rows}x{cols} with a Z node: trace distance {td:.2e}", flush=True)
    assert td < TD_TOL
    # run(): the outcomes dict holds the Z node's expectation p1
    sim.reset(input_state=st[0])
    rho = sim.run(th[0], mode="expectation")
    oracle = ReferenceDM(g, st[0], p_depol=0.01)
    want = oracle.run(th32[0], mode="expectation")
    assert float(_trace_distance(rho, want)) < TD_TOL
    for v, o in oracle.outcomes.items():
        assert abs(sim.outcomes[v] - o) < 1e-5, (v, sim.outcomes[v], o)


def test_dm_impossible_outcome_zero_takes_outcome_one(rng):
    """Input |1> and a first node measured in the XZ plane at pi/2 (a Z
    measurement): outcome 0 has probability 0, the reference takes outcome 1,
    and so does the kernel (outcome reported, state equal to the oracle's)."""
    _torch()
    import mentpy_amd as mp
    from oracle.reference_dm import ReferenceDM
    g = mp.templates.linear_cluster(6)
    first = [v for v in g.measurement_order if v not in g.output_nodes][0]
    g[first] = mp.Ment(np.pi / 2, "XZ")
    th = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes))
    one = np.array([0, 1], dtype=complex)
    sim = mp.PatternSimulator(g, input_state=one, backend="hip-dm")
    rho = sim.run(th)
    oracle = ReferenceDM(g, one)
    want = oracle.run(th.astype(np.float32).astype(np.float64))
    assert oracle.outcomes[first] == 1 and sim.outcomes[first] == 1
    assert np.abs(rho - want).max() < ENTRY_TOL


def test_step_probs_flag_equals_fast_xy_kernel(rng):
    """On an XY pattern with flow every probability is 1/2: the
    per-step-probability kernel and the fast kernels agree."""
    torch = _torch()
    import mentpy_amd as mp
    from mentpy_amd.simulators.hip_simulator import _as_angles, _as_states, dm_forward
    for rows, cols in ((3, 12), (5, 4)):
        g = mp.templates.grid_cluster(rows, cols)
        sim = mp.PatternSimulator(g, backend="hip-dm", p_depol=0.01, p_readout=0.005)
        core = sim.simulator
        a = _as_angles(rng.uniform(0, 2 * np.pi, (40, len(g.trainable_nodes))),
                       core.plan.info["n_angles"], core.device)
        s = _as_states(haar_state(rows, rng)[None], rows, core.device)
        fast = dm_forward(core.plan, a, s, noise=core.noise)
        oc = torch.empty((40, 1, core.plan.info["n_steps"]), dtype=torch.float32, device="cuda")
        slow = dm_forward(core.plan, a, s, noise=core.noise, outcomes=oc)
        err = (fast - slow).abs().max().item()
        print(f"{rows}x{cols}: fast vs per-step-probability kernel {err:.2e}", flush=True)
        assert err < 2e-6
        assert oc.abs().max().item() == 0.0           # every outcome 0


def test_dm_measure_steps_in_mixed_planes(rng):
    """Step-by-step measure() on a mixed-plane pattern: the window after every
    step equals the step-by-step oracle's, outcomes included."""
    _torch()
    import mentpy_amd as mp
    from oracle.reference_dm import ReferenceDM
    g = _mixed(3, 4, rng)
    psi = haar_state(3, rng)
    sim = mp.PatternSimulator(g, input_state=psi, backend="hip-dm", p_depol=0.02)
    ref = ReferenceDM(g, psi, p_depol=0.02)
    worst = 0.0
    for node in sim.simulator.schedule_measure:
        m = g[node]
        th = float(np.float32(rng.uniform(0, 2 * np.pi))) if m.angle is None else m.angle
        got, out = sim.measure(th)
        want, wout = ref.measure(th)
        assert out == wout
        worst = max(worst, float(np.abs(got - want).max()))
    print(f"mixed-plane 3x4 measure(): max window err {worst:.2e}", flush=True)
    assert worst < 2e-6
