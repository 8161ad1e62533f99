"""dev_mode scheduling (CPU).

mentpy's development scheduler (`np_simulator_sv.py:130-162, 176-197,
254-276`, the same in `np_simulator_dm.py`) keeps the window as a list of
nodes and measures, at each step, the first listed node whose first later
neighbour on its wire is listed too (or that has none); nodes still enter the
window in schedule order.  Checked here without a GPU:

* the order the hip backends compute (``dev_mode_order``) is the one the
  step-by-step oracle (``oracle/reference_sv.py`` with ``dev_mode``) follows;
* the oracle is pinned by the full-register simulation of the graph without
  the edges the window drops: projections on distinct qubits commute with
  the CZs of other qubits, so only the dropped edges change the result; they
  are computed here from schedule positions, independently of both;
* plans compiled with such an order (``measure_order``, off the slot
  rotation), executed by the numpy model of the kernel arithmetic, reproduce
  the oracle, and so do the reference windows after every step;
* grids with their rows as wires are measured in schedule order: the plan is
  the one compiled without dev_mode (the fast register kernels run it);
* missing wires and a node on no wire raise ``TypeError`` as in mentpy; the
  C compiler rejects a measure_order that measures a node before it entered
  the window, repeats a node or names an unmeasured one.
"""
from types import SimpleNamespace

import networkx as nx
import numpy as np
import pytest

import mentpy_amd as mp
from conftest import haar_state
from oracle.brute import _node_angle, brute_dm, brute_graph_state
from oracle.reference_dm import ReferenceDM
from oracle.reference_sv import ReferenceSV
from test_plan import _slot_model


def _rows(r, c):
    return [[i * c + j for j in range(c)] for i in range(r)]


def _cases(n=40, seed=5):
    """Grid patterns with schedules (inputs, the other measured nodes row-major
    or shuffled, outputs) and windows for which dev_mode runs and leaves
    schedule order."""
    from mentpy_amd.simulators.hip_simulator import dev_mode_order
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        r, c = int(rng.integers(2, 4)), int(rng.integers(3, 6))
        g = mp.templates.grid_cluster(r, c)
        outs, ins = list(g.output_nodes), list(g.input_nodes)
        rest = [v for v in sorted(g.graph.nodes) if v not in outs and v not in ins]
        if rng.random() >= 0.5:
            rest = [int(v) for v in rng.permutation(rest)]
        sched = ins + rest + outs            # inputs first, as the reference requires
        w = int(rng.integers(r, min(7, len(sched) - len(outs)) + 1))
        try:
            order = dev_mode_order(g, sched, [v for v in sched if v not in outs], w, _rows(r, c))
        except ValueError:
            continue
        if order != [v for v in sched if v not in outs]:
            out.append((r, c, sched, w, order))
    return out


CASES = _cases()


@pytest.fixture(scope="module")
def native_lib():
    from mentpy_amd import _native
    if not _native.library_available():
        from mentpy_amd.csrc.build import build
        build()
    return _native


def _kept_edges(g, sched, w, order):
    """Edges the window applies: (a, b) with a before b in the schedule is
    applied iff a is still unmeasured when b enters (b in the first window,
    or after pos(b) - w + 1 measurements)."""
    pos = {v: i for i, v in enumerate(sched)}
    step = {v: k for k, v in enumerate(order)}
    kept = []
    for a, b in g.graph.edges():
        if pos[a] > pos[b]:
            a, b = b, a
        if pos[b] < w or step.get(a, len(sched)) >= pos[b] - w + 1:
            kept.append((a, b))
    return kept


def _brute_kept(g, kept, th, psi, measured):
    G2 = nx.Graph()
    G2.add_nodes_from(g.graph.nodes())
    G2.add_edges_from(kept)
    full = brute_graph_state(SimpleNamespace(graph=G2, input_nodes=g.input_nodes), psi)
    live = list(G2.nodes())
    for v in measured:
        ax = live.index(v)
        full = np.tensordot(np.array([1.0, np.exp(-1j * _node_angle(g, v, th))]), full,
                            axes=([0], [ax]))
        live.pop(ax)
        full /= np.linalg.norm(full)
    return np.transpose(full, [live.index(v) for v in g.output_nodes]).reshape(-1)


def _fid(a, b):
    return abs(np.vdot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b))


def test_cases_leave_schedule_order():
    assert len(CASES) == 40


@pytest.mark.parametrize("case", range(0, 40, 3))
def test_oracle_dev_mode_equals_brute_force_without_dropped_edges(case, rng):
    r, c, sched, w, order = CASES[case]
    g = mp.templates.grid_cluster(r, c)
    th = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes))
    psi = haar_state(r, rng)
    ref = ReferenceSV(g, psi, window_size=w, schedule=sched, dev_mode=True, wires=_rows(r, c))
    out = ref.run(th, output_form="sv")
    assert list(ref.outcomes) == order                  # the order the backends compute
    kept = _kept_edges(g, sched, w, order)
    want = _brute_kept(g, kept, th, psi, order)
    assert _fid(out, want) > 1 - 1e-10
    # planted mistake: applying every edge (as if nothing were dropped) differs
    if len(kept) < g.graph.number_of_edges():
        assert _fid(out, _brute_kept(g, list(g.graph.edges()), th, psi, order)) < 1 - 1e-6


@pytest.mark.parametrize("case", range(40))
def test_dev_mode_plan_tables_reproduce_oracle(case, rng):
    from mentpy_amd.simulators import compile_pattern
    r, c, sched, w, order = CASES[case]
    g = mp.templates.grid_cluster(r, c)
    plan, _, sm, _ = compile_pattern(g, schedule=sched, window_size=w, dev_mode=True,
                                     wires=_rows(r, c))
    assert plan.measure_order == order
    assert plan.info["rotating"] == 0 and plan.info["periodic"] == 0
    assert plan.info["can_sample"] == 0
    th = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes))
    psi = haar_state(r, rng)
    got = _slot_model(plan, th, psi)
    want = ReferenceSV(g, psi, window_size=w, schedule=sched, dev_mode=True,
                       wires=_rows(r, c)).run(th, output_form="sv")
    assert _fid(got, want) > 1 - 1e-9
    kept = _kept_edges(g, sched, w, order)
    assert plan.info["n_dropped_edges"] == g.graph.number_of_edges() - len(kept)


@pytest.mark.parametrize("rows,cols", [(2, 5), (3, 4), (4, 6)])
def test_grid_rows_as_wires_keep_schedule_order(rows, cols, native_lib):
    from mentpy_amd.simulators import compile_pattern
    g = mp.templates.grid_cluster(rows, cols)
    dev, *_ = compile_pattern(g, dev_mode=True, wires=_rows(rows, cols))
    plain, *_ = compile_pattern(g)
    assert dev.measure_order == [v for v in g.measurement_order if v not in g.output_nodes]
    assert dev.info == plain.info and dev.info["rotating"] == 1
    for a, b in zip(dev.tables(), plain.tables()):
        assert np.array_equal(a, b)


def test_dev_mode_argument_errors(native_lib):
    g = mp.templates.grid_cluster(2, 4)
    with pytest.raises(TypeError):
        mp.PatternSimulator(g, backend="hip-sv", dev_mode=True)           # no wires
    with pytest.raises(TypeError):
        mp.PatternSimulator(g, backend="hip-sv", dev_mode=True, wires=[[0, 1, 2, 3]])
    sim = mp.PatternSimulator(g, backend="hip-sv", dev_mode=True, wires=_rows(2, 4))
    assert sim.simulator.current_simulated_nodes() == g.measurement_order[:3]


def test_measure_order_validation(native_lib, rng):
    from mentpy_amd import _native as nat
    g = mp.templates.grid_cluster(2, 3)          # schedule [0, 3, 1, 4, 2, 5], window 3
    sched = list(g.measurement_order)
    n = len(sched)
    kw = dict(n_nodes=n, edges=np.array(list(g.graph.edges())), schedule=sched, n_measure=4,
              window=3, input_nodes=[0, 3], output_nodes=[2, 5],
              angle_index=[0, 1, -1, 2, 3, -1], fixed_angle=np.zeros(n), n_angles=4)
    ok = nat.NativePlan(**kw, measure_order=[3, 0, 1, 4])
    assert ok.info["rotating"] == 0 and ok.info["lag"] == 1
    # node 1 (position 2) measured at step 0: it entered with the first window,
    # but it is the node a lag-1 plan would append at step 0 -> lag 0
    lag0 = nat.NativePlan(**kw, measure_order=[1, 0, 3, 4])
    assert lag0.info["lag"] == 0 and lag0.info["n_slots"] == 3
    # both orders through the table model against the brute force without
    # the edges the window drops
    for plan, order in ((ok, [3, 0, 1, 4]), (lag0, [1, 0, 3, 4])):
        th = rng.uniform(0, 2 * np.pi, 4)
        psi = haar_state(2, rng)
        kept = _kept_edges(g, sched, 3, order)
        assert plan.info["n_dropped_edges"] == g.graph.number_of_edges() - len(kept)
        assert _fid(_slot_model(plan, th, psi), _brute_kept(g, kept, th, psi, order)) > 1 - 1e-10
    for bad in ([4, 0, 3, 1],      # position 3 has not entered at step 0
                [0, 0, 3, 1],      # repeated
                [0, 3, 1, 2]):     # node 2 is an output
        with pytest.raises(ValueError):
            nat.NativePlan(**kw, measure_order=bad)


def test_dev_mode_window_lists_follow_the_oracle(native_lib, rng):
    """The window's node list after every step: the backend's
    current_simulated_nodes() (which the window kernels' map follows) equals
    the oracle's list, which is updated as the reference updates it."""
    for case in (0, 7, 21):
        r, c, sched, w, order = CASES[case]
        g = mp.templates.grid_cluster(r, c)
        sim = mp.PatternSimulator(g, backend="hip-sv", schedule=sched, window_size=w,
                                  dev_mode=True, wires=_rows(r, c)).simulator
        ref = ReferenceSV(g, haar_state(r, rng), window_size=w, schedule=sched, dev_mode=True,
                          wires=_rows(r, c))
        for k, v in enumerate(order):
            assert ref._dev_pick()[0] == v
            ref.measure(float(rng.uniform(0, 2 * np.pi)))
            sim.current_measurement = k + 1
            assert sim.current_simulated_nodes() == ref.current_simulated_nodes()


# cases with no dropped edge on at most 10 nodes (the brute-force density
# matrix has 4^nodes entries), and a few others
DM_CASES = [c for c in range(40) if CASES[c][0] * CASES[c][1] <= 10
            and len(_kept_edges(mp.templates.grid_cluster(CASES[c][0], CASES[c][1]), CASES[c][2],
                                CASES[c][3], CASES[c][4]))
            == mp.templates.grid_cluster(CASES[c][0], CASES[c][1]).graph.number_of_edges()][:6]
N_BRUTE = len(DM_CASES)
DM_CASES += [0, 13, 29]


@pytest.mark.parametrize("case", DM_CASES)
def test_dm_oracle_dev_mode(case, rng):
    """The density-matrix oracle in dev_mode: at p = 0 the SV oracle's
    density matrix; with noise and no dropped edge, the brute-force Kraus
    simulation in gate order (channels and projections on distinct qubits
    commute, so the measurement order does not matter)."""
    r, c, sched, w, order = CASES[case]
    g = mp.templates.grid_cluster(r, c)
    th = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes))
    psi = haar_state(r, rng)
    kw = dict(window_size=w, schedule=sched, dev_mode=True, wires=_rows(r, c))
    sv = ReferenceSV(g, psi, **kw).run(th, output_form="sv")
    dm0 = ReferenceDM(g, psi, **kw)
    rho0 = dm0.run(th)
    assert list(dm0.outcomes) == order
    assert np.abs(rho0 - np.outer(sv, sv.conj())).max() < 1e-10
    rho = ReferenceDM(g, psi, p_depol=0.02, p_readout=0.01, **kw).run(th)
    # (the full-register density matrix is 4^nodes entries: small grids only)
    if case in DM_CASES[:N_BRUTE]:
        want = brute_dm(g, th, psi, p_depol=0.02, p_readout=0.01)
        assert np.abs(rho - want).max() < 1e-10
