"""Host side (CPU only): pattern model, plan compiler, C ABI.

* The product's schedule / trainable order equals the oracle's restatement of
  mentpy's causal flow and calculate_order, and the grid closed form.
* The plan tables, executed by a numpy model of the kernel arithmetic
  (slot layout, pair coefficients, growth exponent, output gather), reproduce
  the oracle: this checks every table the GPU reads, without a GPU.
* The native library loads and exports every entry point of include/mbqc_hip.h,
  and its validation errors map to mentpy's exception types.
"""
import re
from pathlib import Path

import numpy as np
import pytest

import mentpy_amd as mp
from conftest import haar_state
from oracle.pattern import cflow_layers, grid_order_closed_form, measurement_order
from oracle.reference_sv import ReferenceSV
from oracle.vectorized import sv_batch

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def native():
    from mentpy_amd import _native
    if not _native.library_available():
        from mentpy_amd.csrc.build import build
        build()
    return _native


@pytest.mark.parametrize("maker", [
    lambda: mp.templates.grid_cluster(4, 20), lambda: mp.templates.grid_cluster(3, 12),
    lambda: mp.templates.grid_cluster(2, 5), lambda: mp.templates.linear_cluster(9),
    lambda: mp.templates.many_wires([3, 4, 2])])
def test_schedule_matches_oracle_restatement(maker):
    g = maker()
    flow, layer = cflow_layers(g.graph, g.input_nodes, g.output_nodes)
    assert flow is not None
    assert g.measurement_order == measurement_order(g.graph, g.input_nodes, layer)
    assert g.trainable_nodes == [v for v in g.graph.nodes() if v not in g.output_nodes]
    for v, f in flow.items():
        assert g.flow(v) == f


@pytest.mark.parametrize("n,m", [(4, 20), (2, 3), (5, 40), (3, 12)])
def test_grid_schedule_closed_form(n, m):
    g = mp.templates.grid_cluster(n, m)
    assert g.measurement_order == grid_order_closed_form(n, m)
    assert g.input_nodes == [i * m for i in range(n)]
    assert g.output_nodes == [i * m + m - 1 for i in range(n)]


def _slot_model(plan, angles, psi):
    """numpy complex128 model of the kernel arithmetic driven by the plan tables."""
    st, in_index, sign, out_pos = plan.tables()
    sv_res, _, e = plan.rescale()
    ns = plan.info["n_slots"]
    H = 1 << ns
    t = np.where(sign.astype(bool), -1, 1) * np.asarray(psi)[in_index]
    n = plan.info["n_steps"]
    for k in range(n):
        if k % ns == 0 and sv_res[k // ns]:
            t = t * 2.0 ** -32
        ai, fl, nm, fx = st[k]
        th = angles[int(ai)] if ai >= 0 else fx
        b = np.exp(-1j * th)
        q = (int(fl) >> 8) & 0xFF
        app, sneg = int(fl) & 1, int(fl) & 2
        new = t.copy()
        for j in range(H):
            if j & (1 << q):
                continue
            x, y = t[j], t[j | (1 << q)]
            u0 = x + b * y
            s = -1.0 if sneg else 1.0
            sg = -1.0 if bin(j & int(nm)).count("1") % 2 else 1.0
            new[j] = u0
            new[j | (1 << q)] = sg * (x + s * b * y) if app else 0.0
        t = new
    scale = 2.0 ** (-e[0] / 2)
    out = np.zeros(1 << plan.info["n_outputs"], complex)
    for j in range(H):
        if out_pos[j] >= 0:
            out[out_pos[j]] = t[j] * scale
    return out


@pytest.mark.parametrize("rows,cols,w", [(1, 5, 1), (2, 4, 1), (3, 5, 1), (4, 20, 1), (2, 5, 5),
                                         (2, 4, 4), (3, 4, 6)])
def test_plan_tables_reproduce_oracle(native, rows, cols, w, rng):
    from mentpy_amd.simulators import compile_pattern
    g = mp.templates.grid_cluster(rows, cols) if rows > 1 else mp.templates.linear_cluster(cols)
    if rows > 1:
        g[g.outputc[1]] = mp.Ment("Y")
    plan, *_ = compile_pattern(g, window_size=w)
    th = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes))
    psi = haar_state(rows, rng)
    got = _slot_model(plan, th, psi)
    want = sv_batch(g, th[None], psi[None], window_size=w, phase="raw")[0, 0]
    # fixed angles are stored as float32 in the step table (pi/2 to ~4e-8)
    assert np.abs(got - want).max() < 1e-6
    assert plan.info["n_dropped_edges"] == 0
    assert plan.info["min_window"] <= plan.info["window"]


def test_too_small_window_drops_edges_like_the_reference(native, rng):
    from mentpy_amd.simulators import compile_pattern
    g = mp.templates.grid_cluster(3, 4)
    plan, *_ = compile_pattern(g, window_size=3)
    assert plan.info["min_window"] == 4
    assert plan.info["n_dropped_edges"] > 0
    th = rng.uniform(0, 2 * np.pi, len(g.trainable_nodes))
    psi = haar_state(3, rng)
    want = ReferenceSV(g, psi, window_size=3).run(th, output_form="sv")
    got = _slot_model(plan, th, psi)
    ov = np.vdot(want, got)
    assert abs(abs(ov) / np.linalg.norm(got) - 1) < 1e-10


def test_library_exports_every_header_symbol(native):
    header = (ROOT / "include" / "mbqc_hip.h").read_text()
    names = set(re.findall(r"\b(mbqc_\w+)\s*\(", header))
    assert names, "no entry points found in the header"
    L = native.lib()
    for nm in sorted(names):
        assert hasattr(L, nm), f"{nm} not exported"
    assert set(native.EXPORTED_SYMBOLS) == names
    assert L.mbqc_abi_version() == 3


def test_validation_errors_map_to_reference_exceptions(native):
    from mentpy_amd.simulators import compile_pattern
    g = mp.templates.grid_cluster(2, 3)
    with pytest.raises(ValueError, match="Window size"):
        compile_pattern(g, window_size=10)
    with pytest.raises(ValueError, match="Input state has"):
        compile_pattern(mp.templates.grid_cluster(3, 4), window_size=2)
    with pytest.raises(ValueError):
        mp.PatternSimulator(g, backend="no-such-backend")
    with pytest.raises(TypeError):
        mp.PatternSimulator(g, input_state=[1, 0, 0, 0], backend="hip-sv")
    with pytest.raises(NotImplementedError):
        mp.PatternSimulator(g, backend="hip-dm", force0=False)
    g2 = mp.templates.grid_cluster(2, 3)
    g2[1] = mp.Ment("XZ")
    with pytest.raises(ValueError, match="only XY plane"):
        mp.PatternSimulator(g2, backend="hip-sv")
    sim = mp.PatternSimulator(g2, backend="hip-dm")       # the DM backend has no plane check
    assert sim.simulator.plan.info["general_planes"] == 1
    L = native.lib()
    rc = L.mbqc_sv_forward_c64(sim.simulator.plan.handle, None, 0, 1, None, 0, None, None, 1, 1, 0,
                               None)
    assert rc == native.MBQC_ERR_VALUE and "only XY plane" in L.mbqc_last_error().decode()
    g3 = mp.templates.grid_cluster(2, 3)
    with pytest.warns(UserWarning):
        g3[1] = mp.Ment((0.1, 0.2), "XYZ")
    with pytest.raises(NotImplementedError, match="pair of angles"):
        mp.PatternSimulator(g3, backend="hip-dm")
    g4 = mp.templates.grid_cluster(2, 3)
    g4[1] = mp.Ment("Z")
    with pytest.raises(NotImplementedError, match="expectation"):
        mp.PatternSimulator(g4, backend="hip-dm").run([0.1] * len(g4.trainable_nodes))
    with pytest.raises(ValueError, match="Number of angles"):
        mp.PatternSimulator(g, backend="hip-sv").run([0.1])


def test_kernel_limits_are_reported(native):
    """Windows beyond the kernels' reach fail with NotImplementedError before
    any device work (the reference has no limit, but no configuration needs more)."""
    from mentpy_amd.simulators import compile_pattern
    plan, *_ = compile_pattern(mp.templates.grid_cluster(8, 9), exclude="quantum_output")
    assert plan.info["n_slots"] == 8
    L = native.lib()
    plan11, *_ = compile_pattern(mp.templates.grid_cluster(11, 3))
    assert plan11.info["n_slots"] == 11
    rc = L.mbqc_dm_forward_c64(plan11.handle, None, 0, 1, None, 0, None, 1, 1, None, 0, None)
    assert rc == native.MBQC_ERR_NOT_IMPLEMENTED          # density matrices: at most 10 slots
    rc = L.mbqc_sv_segmented_forward_c64(plan.handle, None, 0, 1, None, 1, 2, None, 0, None)
    assert rc == native.MBQC_ERR_NOT_IMPLEMENTED          # time segments: at most 5 slots
    assert L.mbqc_sv_segments_auto(plan.handle, 1) == 1
    plan17, *_ = compile_pattern(mp.templates.grid_cluster(17, 3))
    rc = L.mbqc_sv_adjoint_forward_c64(plan17.handle, None, 0, 1, None, 0, None, 1, 1, None, 0, None)
    assert rc == native.MBQC_ERR_NOT_IMPLEMENTED          # adjoint: at most 16 slots
    plan20, *_ = compile_pattern(mp.templates.grid_cluster(20, 3))
    assert plan20.info["n_slots"] == 20                     # statevectors: up to 20 slots
    with pytest.raises(NotImplementedError):
        compile_pattern(mp.templates.grid_cluster(21, 3))


def test_adjoint_rejects_shared_angle_entries(native):
    """An angle-vector entry used by two measured nodes runs in the forward
    kernels, but the adjoint entry points (which write one gradient entry per
    step) reject it before any device work."""
    import numpy as np
    g = mp.templates.grid_cluster(2, 4)
    from mentpy_amd.simulators import compile_pattern
    plan, sched, sched_meas, w = compile_pattern(g)
    nodes = list(g.graph.nodes())
    tr = list(g.trainable_nodes)
    ai = np.full(len(nodes), -1, np.int32)
    for v in sched_meas:
        ai[nodes.index(v)] = tr.index(v) % 3          # three angles, shared
    shared = native.NativePlan(
        n_nodes=len(nodes), edges=[(nodes.index(a), nodes.index(b)) for a, b in g.graph.edges()],
        schedule=[nodes.index(v) for v in sched], n_measure=len(sched_meas), window=w,
        input_nodes=[nodes.index(v) for v in g.input_nodes],
        output_nodes=[nodes.index(v) for v in g.output_nodes], angle_index=ai,
        fixed_angle=np.zeros(len(nodes), np.float32), n_angles=3)
    L = native.lib()
    for p, want in ((shared, native.MBQC_ERR_NOT_IMPLEMENTED),):
        rc = L.mbqc_sv_adjoint_forward_c64(p.handle, None, 0, 1, None, 0, None, 1, 1, None, 0, None)
        assert rc == want
        assert "at most one measured node" in L.mbqc_last_error().decode()
        rc = L.mbqc_sv_adjoint_backward_c64(p.handle, None, 0, 1, None, 1, 1, None, 0, None, None)
        assert rc == want
        rc = L.mbqc_sv_segmented_forward_c64(p.handle, None, 0, 1, None, 1, 2, None, 0, None)
        assert rc == want


def test_circuit_noise_names_map_to_noise_model():
    """The PennyLane backend's circuit_noise names (pennylane_simulator.py:123-136)."""
    nm = mp.NoiseModel.from_circuit_noise("generalized_amplitude_damping", p=0.3, gamma=0.2)
    assert (nm.gad_gamma, nm.gad_p) == (0.2, 0.3)
    assert mp.NoiseModel.from_circuit_noise("phase_flip", p=0.1).phase_flip == 0.1
    assert mp.NoiseModel.from_circuit_noise("amplitude_damping", gamma=0.4).amplitude_damping == 0.4
    with pytest.raises(ValueError, match="Unrecognized circuit noise"):
        mp.NoiseModel.from_circuit_noise("bit_flip", p=0.1)
    with pytest.raises(ValueError):
        mp.NoiseModel(amplitude_damping=1.5)
    g = mp.templates.grid_cluster(2, 3)
    with pytest.raises(ValueError, match="Unrecognized circuit noise"):
        mp.PatternSimulator(g, backend="hip-dm", circuit_noise="bit_flip", p=0.1)
    with pytest.raises(ValueError):
        mp.PatternSimulator(g, backend="hip-dm", circuit_noise="phase_flip", p=0.1, p_depol=0.1)
    sim = mp.PatternSimulator(g, backend="hip-dm", circuit_noise="phase_damping", gamma=0.3)
    assert sim.simulator.noise.phase_damping == 0.3


def test_host_only_abi_functions(native):
    """Host-side entry points that need no GPU: window sizes of the step API,
    the Gram workspace size and the causal-flow error codes."""
    import ctypes as C
    from mentpy_amd.simulators import compile_pattern
    L = native.lib()
    plan, schedule, meas, w = compile_pattern(mp.templates.grid_cluster(3, 5))
    N, n = len(schedule), len(meas)
    for k in range(n + 1):
        assert L.mbqc_plan_window_qubits(plan.handle, k) == min(w, N - k)
    assert L.mbqc_plan_window_qubits(plan.handle, n + 1) == native.MBQC_ERR_VALUE
    assert L.mbqc_plan_window_qubits(plan.handle, -1) == native.MBQC_ERR_VALUE
    assert L.mbqc_gram_workspace_bytes(1024, 256) == 1024 * 16 * 4
    assert L.mbqc_gram_workspace_bytes(-1, 256) == 0
    e = np.array([[0, 1], [1, 2]], np.int32)
    ins, outs = np.array([0], np.int32), np.array([1, 2], np.int32)
    found = C.c_int32(-1)
    rc = L.mbqc_causal_flow(3, 2, C.c_void_p(e.ctypes.data), 1, C.c_void_p(ins.ctypes.data), 2,
                            C.c_void_p(outs.ctypes.data), None, None, None, C.byref(found))
    assert rc == native.MBQC_ERR_VALUE
    assert b"different size" in L.mbqc_last_error()
    # a graph without causal flow: found = 0, no error
    e2 = np.array([[0, 1], [1, 2], [2, 0], [0, 3], [3, 2]], np.int32)
    ins2, outs2 = np.array([0], np.int32), np.array([3], np.int32)
    rc = L.mbqc_causal_flow(4, 5, C.c_void_p(e2.ctypes.data), 1, C.c_void_p(ins2.ctypes.data), 1,
                            C.c_void_p(outs2.ctypes.data), None, None, None, C.byref(found))
    assert rc == native.MBQC_OK
    from mentpy_amd.mbqc.flow import find_cflow
    import networkx as nx
    G = nx.Graph([tuple(x) for x in e2.tolist()])
    assert (found.value == 1) == (find_cflow(G, [0], [3])[0] is not None)


def test_segmented_adjoint_host_sizes(native):
    """Segment count choice and workspace sizes of the time-segmented adjoint
    (host-only): 4x16 grid = 15 blocks of 4 steps, 16 columns per segment."""
    from mentpy_amd.simulators import compile_pattern
    L = native.lib()
    plan = compile_pattern(mp.templates.grid_cluster(4, 16))[0]
    assert L.mbqc_sv_segments_auto(plan.handle, 1024) == 4
    assert L.mbqc_sv_segments_auto(plan.handle, 4096) == 2
    assert L.mbqc_sv_segments_auto(plan.handle, 16384) == 1
    assert L.mbqc_sv_segments_auto(plan.handle, 1) == 15          # capped by the block count
    assert L.mbqc_sv_segmented_workspace_bytes(plan.handle, 1024, 1) == 0
    assert L.mbqc_sv_segmented_workspace_bytes(plan.handle, 1024, 16) == 0
    ws = L.mbqc_sv_segmented_workspace_bytes(plan.handle, 1024, 8)
    # checkpoints (15 blocks x 16 amplitudes x 16 columns per set) dominate
    ckpt = 15 * 16 * 1024 * 16 * 8
    assert ckpt < ws < ckpt + 8 * 4 * 1024 * 16 * 16 * 8
    assert L.mbqc_sv_segmented_forward_c64(plan.handle, None, 0, 1, None, 4, 1, None, 0, None) \
        == native.MBQC_ERR_VALUE
    assert b"n_segments" in L.mbqc_last_error()


def test_coefficient_rows_per_step_or_deduplicated(native):
    """Step k's pair coefficients are row flags >> 16 of the plan's table: one
    row per step (row k) while the table fits 8 KB, each distinct row once
    beyond that (a grid has one per slot)."""
    from mentpy_amd.simulators import compile_pattern
    for rows, cols in [(4, 20), (4, 40), (5, 30), (6, 20), (2, 40)]:
        plan = compile_pattern(mp.templates.grid_cluster(rows, cols))[0]
        st = plan.tables()[0]
        ns, n = plan.info["n_slots"], plan.info["n_steps"]
        idx = np.array([int(x[1]) >> 16 for x in st])
        if n * 4 << ns <= 8192:
            assert np.array_equal(idx, np.arange(n)), (rows, cols)
        else:
            assert idx.max() + 1 <= 2 * ns, (rows, cols, idx.max())
            # a row is shared only by steps with the same measured slot
            for r in set(idx.tolist()):
                assert len({(int(x[1]) >> 8) & 0xFF for x, i in zip(st, idx) if i == r}) == 1


def test_periodic_plans(native):
    """Grids in column order and chains are periodic: every step appends a CZ
    neighbour of the measured qubit and the appended qubit's other CZ
    neighbours repeat every n_slots steps, so the forward kernel fixes the pair
    coefficients at compile time.  Other patterns and wider windows are not."""
    from mentpy_amd.simulators import compile_pattern
    cases = [(mp.templates.linear_cluster(9), None, 1)]
    cases += [(mp.templates.grid_cluster(r, c), None, 1)
              for r, c in [(2, 5), (3, 4), (4, 20), (4, 16), (5, 7), (6, 8)]]
    cases += [(mp.templates.grid_cluster(3, 4), 5, 0),      # window above the minimum
              (mp.templates.grid_cluster(7, 3), None, 0),   # 7 slots: no coefficient table
              (mp.templates.muta(2, 1), None, 0)]
    for g, w, expect in cases:
        plan = compile_pattern(g, window_size=w or 1)[0]
        info = plan.info
        assert info["periodic"] == expect, (len(g.graph), w)
        if expect:
            st = plan.tables()[0]
            ns, n = info["n_slots"], info["n_steps"]
            flags = st[:, 1].astype(np.int64)
            nmask = st[:, 2].astype(np.int64)
            assert n % ns == 0
            assert np.all(flags & 3 == 3)
            assert np.array_equal(nmask, np.tile(nmask[:ns], n // ns))
            # the appended qubit's earlier neighbour sits in the slot appended one step before
            assert nmask[0] == 0 and all(nmask[s] == 1 << (ns - s) for s in range(1, ns))


def test_finalizers_defer_native_destruction(native):
    """A plan held in a reference cycle (simulators and their backends form
    one) dies in the cyclic garbage collector, which can run at any allocation,
    including inside a HIP graph capture, where freeing device memory breaks
    the capture: its finalizer only queues the native handle, and the next plan
    construction (never inside a capture) destroys it."""
    import gc
    from mentpy_amd.simulators import compile_pattern
    native.release_pending()
    g = mp.templates.grid_cluster(2, 3)
    plan, *_ = compile_pattern(g)
    h = plan.handle.value
    holder = {"plan": plan}
    holder["self"] = holder
    del plan, holder
    gc.collect()
    assert ("mbqc_plan_destroy", h) in native._pending
    plan2, *_ = compile_pattern(g)
    assert ("mbqc_plan_destroy", h) not in native._pending
    assert plan2.info["n_slots"] == 2
