"""MI355X backends ``hip-sv`` and ``hip-dm``.

They keep the constructor keywords, ``run``/``reset``/``measure`` semantics,
error types and return forms of the reference's numpy backends
(`mentpy/simulators/np_simulator_sv.py:38-384`,
`mentpy/simulators/np_simulator_dm.py:33-380`): the schedule is the
pattern's measurement order (or ``schedule=``), the window defaults to
``|I|+1`` when a flow exists, every outcome is post-selected to 0, ``run``
returns a numpy density matrix (SV: ``output_form="dm"`` default, or
``"sv"``), and a second ``run`` without ``reset`` raises ``ValueError``.

The arithmetic runs in the HIP kernels of ``libmbqc_hip.so`` in complex64;
``run_batch`` exposes the batched device-resident call (torch tensors in HBM,
no host copies).  There is no CPU path: without a GPU the calls raise.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np

from .. import _native as nat
from ..operators.ment import Ment
from .base_simulator import BaseSimulator

__all__ = ["HipSimulatorSV", "HipSimulatorDM", "compile_pattern", "sv_forward", "sv_sample",
           "decode_outcomes", "dm_forward"]


def _torch():
    import torch
    return torch


# ----------------------------------------------------------------- compile
_PLANE_CODES = {"XY": nat.PLANE_XY, "X": nat.PLANE_XY, "Y": nat.PLANE_XY, "XZ": nat.PLANE_XZ,
                "YZ": nat.PLANE_YZ, "Z": nat.PLANE_Z}


def dev_mode_order(mbqcircuit, schedule, schedule_measure, window: int, wires):
    """Node measured at each step by mentpy's ``dev_mode`` scheduler
    (`np_simulator_sv.py:130-162, 176-197, 254-276`; the same in
    `np_simulator_dm.py`): the window is a list of nodes, initially
    ``schedule[:window]``; each step measures the first listed node whose
    first later neighbour on its wire is listed too (or that has none), drops
    it and appends the next scheduled node.  The choice depends on the graph
    and wires only, so it is made once, on the host."""
    if wires is None:
        raise TypeError("dev_mode needs wires: a list of node lists, one per wire")
    wires = [list(w) for w in wires]
    graph = mbqcircuit.graph

    def wire_of(node):
        for wire in wires:
            if node in wire:
                return wire
        raise TypeError(f"node {node} is on none of the wires")

    measured = set(schedule_measure)
    n_nodes = len(schedule)
    window_nodes = list(schedule[:window])
    order = []
    for cm in range(len(schedule_measure)):
        picked = None
        for node in window_nodes:
            later = [n for n in wire_of(node) if graph.has_edge(node, n) and n > node]
            if not later or later[0] in window_nodes:
                picked = node
                break
        if picked is None:
            raise ValueError("dev_mode: no node of the window can be measured "
                             f"(window {window_nodes})")
        if picked not in measured:
            # the reference would measure an output node here (and fail on its
            # missing measurement)
            raise ValueError(f"dev_mode picks unmeasured node {picked}")
        order.append(picked)
        window_nodes = [v for v in window_nodes if v != picked]
        if cm + 1 + window <= n_nodes:
            window_nodes.append(schedule[cm + window])
    return order


def compile_pattern(mbqcircuit, schedule=None, window_size: int = 1, exclude: str = "output",
                    xy_only_error=ValueError, all_planes: bool = False, dev_mode: bool = False,
                    wires=None):
    """Flatten an ``MBQCircuit`` into the arrays of the C plan and compile it.

    ``exclude`` selects the unmeasured nodes: ``"output"`` (statevector
    backend, `np_simulator_sv.py:62-75`) or ``"quantum_output"`` (density
    matrix backend, `np_simulator_dm.py:48-61`).  With ``all_planes`` the
    XZ, YZ and Z planes are compiled too (density-matrix backend, which has
    no plane check, `np_simulator_dm.py:33-115`); otherwise they raise
    ``xy_only_error`` as the statevector backend does
    (`np_simulator_sv.py:54-59`).  With ``dev_mode`` the nodes are measured
    in ``dev_mode_order``.  Returns ``(NativePlan, schedule, schedule_measure,
    window)``; ``NativePlan.measure_order`` lists the node of every step.
    """
    graph = mbqcircuit.graph
    nodes = list(graph.nodes())
    idx = {v: i for i, v in enumerate(nodes)}
    excl = mbqcircuit.output_nodes if exclude == "output" else mbqcircuit.quantum_output_nodes
    if schedule is not None:
        schedule = list(schedule)
        sched_meas = [i for i in schedule if i not in excl]
    elif mbqcircuit.measurement_order is not None:
        schedule = list(mbqcircuit.measurement_order)
        sched_meas = [i for i in schedule if i not in excl]
        if window_size == 1 and mbqcircuit.flow is not None:
            window_size = len(mbqcircuit.input_nodes) + 1
    else:
        raise ValueError("Schedule must be provided for numpy simulator as the MBQCircuit "
                         "does not have a flow.")
    if len(schedule) != len(nodes) or set(schedule) != set(nodes):
        raise ValueError("schedule must contain every node of the pattern exactly once")
    n_meas = len(sched_meas)
    if schedule[:n_meas] != sched_meas:
        raise NotImplementedError(
            "schedules that place unmeasured nodes before measured ones are not supported")
    trainable = list(mbqcircuit.trainable_nodes)
    tpos = {v: k for k, v in enumerate(trainable)}
    angle_index = np.full(len(nodes), -1, np.int32)
    fixed = np.zeros(len(nodes), np.float32)
    plane = np.zeros(len(nodes), np.int32)
    for v in sched_meas:
        m = mbqcircuit[v]
        if not isinstance(m, Ment):
            raise ValueError(f"Ment has no matrix representation at qubit {v}")
        if m.plane not in ("X", "Y", "XY"):
            if not all_planes:
                raise xy_only_error(
                    f"Node {v} has plane {m.plane}, but only XY plane is supported.")
            if m.plane == "XYZ":
                raise NotImplementedError(
                    f"Node {v} is measured in the XYZ plane, which takes a pair of angles per node; "
                    "the angle vectors of the hip backends carry one angle per node")
        plane[idx[v]] = _PLANE_CODES[m.plane]
        if v in tpos:
            angle_index[idx[v]] = tpos[v]
        elif m.plane in ("X", "Y", "XY"):
            fixed[idx[v]] = m.xy_angle()
        elif m.plane != "Z":
            fixed[idx[v]] = float(m.angle)
    edges = np.array([(idx[a], idx[b]) for a, b in graph.edges()], np.int32).reshape(-1, 2)
    outputs = [v for v in schedule[n_meas:]]
    target_out = list(mbqcircuit.output_nodes if exclude == "output"
                      else mbqcircuit.quantum_output_nodes)
    if set(target_out) != set(outputs):
        raise ValueError("output nodes must be the unmeasured nodes of the schedule")
    flow = None
    if mbqcircuit.flow is not None:
        # flow successors of the measured nodes (-1 elsewhere): the byproduct
        # corrections of sampled outcomes (pennylane_simulator.py:146-153)
        flow = np.full(len(nodes), -1, np.int32)
        try:
            for v in sched_meas:
                f = mbqcircuit.flow(v)
                flow[idx[v]] = idx[f] if f in idx else -1
        except (KeyError, IndexError, TypeError, ValueError):
            flow = None
    order = list(sched_meas)
    if dev_mode:
        order = dev_mode_order(mbqcircuit, schedule, sched_meas, int(window_size), wires)
    plan = nat.NativePlan(
        n_nodes=len(nodes), edges=edges, schedule=[idx[v] for v in schedule], n_measure=n_meas,
        window=int(window_size), input_nodes=[idx[v] for v in mbqcircuit.input_nodes],
        output_nodes=[idx[v] for v in target_out], angle_index=angle_index, fixed_angle=fixed,
        n_angles=len(trainable), flow=flow, plane=plane if plane.any() else None,
        measure_order=None if order == sched_meas else [idx[v] for v in order])
    plan.measure_order = order
    return plan, schedule, sched_meas, int(window_size)


# ----------------------------------------------------------------- launches
def _stream_ptr(device):
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _as_angles(angles, n_angles: int, device):
    """float32 device view ``[B, n]``; any strides are kept (an angle-major
    ``[n, B]`` tensor passed as ``.t()`` gives coalesced loads)."""
    torch = _torch()
    a = torch.as_tensor(angles, dtype=torch.float32, device=device)
    if a.dim() == 1:
        a = a.unsqueeze(0)
    if a.dim() != 2 or a.shape[1] != n_angles:
        raise ValueError(f"angles must have shape [B, {n_angles}], got {tuple(a.shape)}")
    if a.stride(1) == 0 and n_angles > 1:
        a = a.contiguous()
    return a


def _angle_strides(a):
    return (a.stride(0) if a.shape[0] > 1 else 0), max(a.stride(1), 1)


def _as_states(states, n_qubits: int, device):
    torch = _torch()
    s = torch.as_tensor(states, device=device)
    if not torch.is_complex(s):
        s = s.to(torch.float32)
    s = s.to(torch.complex64)
    if s.dim() == 1:
        s = s.unsqueeze(0)
    if s.dim() != 2 or s.shape[1] != (1 << n_qubits):
        raise ValueError(f"input states must have shape [S, {1 << n_qubits}], got {tuple(s.shape)}")
    return s.contiguous()


def _require_gpu(device):
    torch = _torch()
    if device.type != "cuda" or not torch.cuda.is_available():
        raise RuntimeError("hip backends need a ROCm GPU device (device='cuda')")


def sv_forward(plan: nat.NativePlan, angles, states, *, ref_phase: bool = True,
               renormalize: bool = True, want_phase: bool = False, out=None,
               table_steps: bool = False):
    """Batched statevector forward: ``angles [B, n]`` x ``states [S, 2^|I|]`` ->
    ``[B, S, 2^|O|]`` complex64 on the device (and the per-simulation global
    phase ``[B, S]`` if ``want_phase``).  ``table_steps`` makes a periodic
    plan (grids, chains) read its pair coefficients from the table like any
    other plan; both give identical results."""
    torch = _torch()
    device = states.device
    _require_gpu(device)
    info = plan.info
    B, S = angles.shape[0], states.shape[0]
    if out is None:
        out = torch.empty((B, S, 1 << info["n_outputs"]), dtype=torch.complex64, device=device)
    phase = torch.empty((B, S), dtype=torch.complex64, device=device) if want_phase else None
    flags = ((nat.FLAG_REF_PHASE if ref_phase else 0) | (nat.FLAG_RENORMALIZE if renormalize else 0)
             | (nat.FLAG_WRITE_PHASE if want_phase else 0)
             | (nat.FLAG_TABLE_STEPS if table_steps else 0))
    sb, sa = _angle_strides(angles)
    with torch.cuda.device(device):
        nat.check(nat.lib().mbqc_sv_forward_c64(
            plan.handle, C.c_void_p(angles.data_ptr()), sb, sa,
            C.c_void_p(states.data_ptr()), states.shape[1], C.c_void_p(out.data_ptr()),
            C.c_void_p(phase.data_ptr() if phase is not None else 0), B, S, flags,
            _stream_ptr(device)))
    return (out, phase) if want_phase else out


def sv_sample(plan: nat.NativePlan, angles, states, seed: int, *, out=None, outcomes=None):
    """Sampled-outcome statevector forward with flow corrections.

    Returns ``(out [B, S, 2^|O|], rows [B, S, W + 2])``: the corrected output
    (equal to the post-selected output up to a global phase) and per
    simulation the outcome bit words (W = ceil(n_steps / 32), bit k % 32 of
    word k // 32 = outcome of the k-th measured node) followed by the X and Z
    frame applied to the output qubits.
    """
    torch = _torch()
    device = states.device
    _require_gpu(device)
    info = plan.info
    B, S = angles.shape[0], states.shape[0]
    W = (info["n_steps"] + 31) // 32
    if out is None:
        out = torch.empty((B, S, 1 << info["n_outputs"]), dtype=torch.complex64, device=device)
    if outcomes is None:
        outcomes = torch.empty((B, S, W + 2), dtype=torch.int32, device=device)
    sb, sa = _angle_strides(angles)
    with torch.cuda.device(device):
        nat.check(nat.lib().mbqc_sv_sample_c64(
            plan.handle, C.c_void_p(angles.data_ptr()), sb, sa,
            C.c_void_p(states.data_ptr()), states.shape[1], C.c_void_p(out.data_ptr()),
            C.c_void_p(outcomes.data_ptr()), B, S, C.c_uint64(int(seed) & (2 ** 64 - 1)),
            _stream_ptr(device)))
    return out, outcomes


def decode_outcomes(rows, n_steps: int):
    """Outcome words ``[..., W (+2)]`` -> bits ``[..., n_steps]`` (numpy uint8)."""
    r = rows.cpu().numpy() if hasattr(rows, "cpu") else np.asarray(rows)
    w = r[..., : (n_steps + 31) // 32].astype(np.int64) & 0xFFFFFFFF
    k = np.arange(n_steps)
    return ((w[..., k // 32] >> (k % 32)) & 1).astype(np.uint8)


def dm_forward(plan: nat.NativePlan, angles, states, *, p_depol: float = 0.0,
               p_readout: float = 0.0, noise=None, renormalize: bool = True, out=None,
               table_steps: bool = False, step_probs: bool = False, outcomes=None):
    """Batched density-matrix forward -> ``[B, S, 2^|O|, 2^|O|]`` complex64.
    ``noise`` (a ``NoiseModel``) replaces ``p_depol`` / ``p_readout``;
    ``table_steps`` makes a periodic plan read its step flags from the tables
    (same results).  ``step_probs`` evaluates every step's outcome
    probability (always done for plans with XZ / YZ / Z nodes); ``outcomes``
    (a float32 ``[B, S, n_steps]`` device tensor) also receives the outcomes."""
    torch = _torch()
    device = states.device
    _require_gpu(device)
    info = plan.info
    B, S = angles.shape[0], states.shape[0]
    d = 1 << info["n_outputs"]
    if out is None:
        out = torch.empty((B, S, d, d), dtype=torch.complex64, device=device)
    if noise is None:
        noise = nat.Noise(p_depol=float(p_depol), p_readout=float(p_readout))
    elif not isinstance(noise, nat.Noise):
        noise = noise.native()
    flags = ((nat.FLAG_RENORMALIZE if renormalize else 0)
             | (nat.FLAG_TABLE_STEPS if table_steps else 0)
             | (nat.FLAG_STEP_PROBS if step_probs else 0))
    sb, sa = _angle_strides(angles)
    with torch.cuda.device(device):
        if outcomes is not None:
            if tuple(outcomes.shape) != (B, S, info["n_steps"]) or outcomes.dtype != torch.float32:
                raise ValueError(f"outcomes must be a float32 tensor of shape {(B, S, info['n_steps'])}")
            nat.check(nat.lib().mbqc_dm_forward_outcomes_c64(
                plan.handle, C.c_void_p(angles.data_ptr()), sb, sa,
                C.c_void_p(states.data_ptr()), states.shape[1], C.c_void_p(out.data_ptr()),
                C.c_void_p(outcomes.data_ptr()), B, S, C.byref(noise), flags, _stream_ptr(device)))
        else:
            nat.check(nat.lib().mbqc_dm_forward_c64(
                plan.handle, C.c_void_p(angles.data_ptr()), sb, sa,
                C.c_void_p(states.data_ptr()), states.shape[1], C.c_void_p(out.data_ptr()), B, S,
                C.byref(noise), flags, _stream_ptr(device)))
    return out


# ----------------------------------------------------------------- backends
class _HipBackend(BaseSimulator):
    _exclude = "output"

    def __init__(self, mbqcircuit, input_state=None, **kwargs):
        super().__init__(mbqcircuit, input_state)
        self.window_size = kwargs.pop("window_size", 1)
        self.schedule = kwargs.pop("schedule", None)
        self.force0 = kwargs.pop("force0", True)
        seed = kwargs.pop("seed", None)
        self.dev_mode = kwargs.pop("dev_mode", False)
        self.wires = kwargs.pop("wires", None)
        self._device_arg = kwargs.pop("device", "cuda")
        self._backend_kwargs(kwargs)
        if kwargs:
            raise TypeError(f"unexpected keyword arguments {sorted(kwargs)}")
        if not self.force0 and not self._can_sample:
            raise NotImplementedError("Numpy simulator does not support force0=False.")
        self._check_planes(mbqcircuit)
        plan, schedule, sched_meas, window = compile_pattern(
            mbqcircuit, self.schedule, self.window_size, self._exclude, self._plane_error,
            all_planes=self._all_planes, dev_mode=self.dev_mode, wires=self.wires)
        self.plan = plan
        # node measured at each step (schedule_measure unless dev_mode reorders)
        self.measure_order = plan.measure_order
        torch = _torch()
        dev = self.device
        if dev.type == "cuda" and torch.cuda.is_available():
            # tables on the backend's device now, so no launch allocates
            with torch.cuda.device(dev):
                plan.upload()
        self.schedule = schedule
        self.schedule_measure = sched_meas
        self.window_size = window
        if not self.force0 and not plan.info["can_sample"]:
            raise NotImplementedError(
                "force0=False needs a pattern with a causal flow, no dropped CZ edges and a "
                "live window of at most 20 qubits")
        self._rng = np.random.default_rng(seed)
        self.last_outcomes = None
        self._step_angles = []
        n_in = len(mbqcircuit.input_nodes)
        if input_state is None:
            input_state = np.ones(1)
            for _ in range(n_in):
                input_state = np.kron(input_state, np.array([1, 1]) / np.sqrt(2))
        self._set_input(input_state)
        self.current_measurement = 0
        self.outcomes = {}

    # hooks
    _plane_error = ValueError
    _all_planes = False
    _can_sample = False

    def _backend_kwargs(self, kwargs):
        pass

    def _check_planes(self, mbqcircuit):
        pass

    @property
    def device(self):
        torch = _torch()
        return torch.device(self._device_arg)

    def _set_input(self, state):
        n_in = len(self.mbqcircuit.input_nodes)
        arr = state.detach().cpu().numpy() if hasattr(state, "detach") else np.asarray(state)
        arr = arr.reshape(-1)
        if arr.shape[0] != (1 << n_in):
            raise ValueError(f"Input state has {int(math.log2(max(arr.shape[0], 1)))} qubits, "
                             f"but the pattern has {n_in} inputs.")
        self.input_state = state
        self._input_np = arr.astype(np.complex128)

    def reset(self, input_state=None):
        self.current_measurement = 0
        self._step_angles = []
        if input_state is not None:
            self._set_input(input_state)
        self.outcomes = {}

    def measure(self, angle: float, **kwargs):
        """Measure the next scheduled node at ``angle`` (post-selected outcome 0)
        and return ``(qstate, outcome)``, the reference's window state after the
        step (`np_simulator_sv.py:164-225`, `np_simulator_dm.py:151-216`).

        The window is recomputed on the GPU from the input through all steps so
        far (one launch per call).
        """
        if self.current_measurement >= len(self.schedule_measure):
            raise ValueError("No more measurements to be done.")
        if not self.force0:
            raise NotImplementedError("step-by-step measure() draws no outcomes; use run()")
        node = self.measure_order[self.current_measurement]
        ment = self.mbqcircuit[node]
        ment.matrix(angle)                 # the reference's angle checks (ment.py:218-229)
        self._check_step_mode(ment, kwargs.get("mode", "sample"))
        if ment.plane == "X":
            theta = 0.0
        elif ment.plane == "Y":
            theta = np.pi / 2
        elif ment.plane == "Z":
            theta = 0.0
        else:
            theta = float(angle if angle is not None else ment.angle)
        self._step_angles.append(theta)
        self.current_measurement += 1
        self.qstate, outcome = self._window(len(self._step_angles))
        self.outcomes[node] = outcome
        return self.qstate, outcome

    def _check_step_mode(self, ment, mode):
        pass

    def _window_args(self, k):
        torch = _torch()
        dev = self.device
        _require_gpu(dev)
        a = torch.tensor(self._step_angles if self._step_angles else [0.0], dtype=torch.float32,
                         device=dev)
        s = _as_states(self._input_np, self.plan.info["n_inputs"], dev)
        W = int(nat.lib().mbqc_plan_window_qubits(self.plan.handle, k))
        if W < 0:
            nat.check(W)
        return a, s, W

    def _check_run(self, angles, input_state):
        if input_state is not None:
            self.reset(input_state=input_state)
        if len(angles) != len(self.mbqcircuit.trainable_nodes):
            raise ValueError(f"Number of angles ({len(angles)}) does not match number of "
                             f"trainable nodes ({len(self.mbqcircuit.trainable_nodes)}).")
        # mentpy's run() measures every scheduled node through measure(), so
        # after any earlier measurement it runs out of nodes and raises
        if self.current_measurement > 0 and self.schedule_measure:
            raise ValueError("No more measurements to be done.")

    def _finish_run(self, bits=None):
        self.current_measurement = len(self.schedule_measure)
        if bits is None:
            self.outcomes = {v: 0 for v in self.measure_order}
        else:
            self.outcomes = {v: int(b) for v, b in zip(self.measure_order, bits)}

    def current_simulated_nodes(self):
        """Nodes of the reference's window, oldest first
        (`np_simulator_sv.py:130-136`): ``schedule[k:k+w]`` after k steps, or in
        dev_mode the entered nodes not yet measured."""
        k, w = self.current_measurement, self.window_size
        if not self.dev_mode:
            return self.schedule[k:k + w]
        done = set(self.measure_order[:k])
        return [v for v in self.schedule[:k + w] if v not in done]

    def current_number_simulated_nodes(self) -> int:
        return min(self.window_size, len(self.mbqcircuit) - self.current_measurement)

    def _next_seed(self) -> int:
        return int(self._rng.integers(0, 2 ** 63, dtype=np.int64))

    def _device_inputs(self, angles):
        torch = _torch()
        dev = self.device
        _require_gpu(dev)
        a = _as_angles(np.asarray(angles, dtype=np.float64).reshape(1, -1),
                       self.plan.info["n_angles"], dev)
        s = _as_states(self._input_np, self.plan.info["n_inputs"], dev)
        return a, s


class HipSimulatorSV(_HipBackend):
    """Statevector backend (complex64 on the GPU).

    Extra keywords: ``device`` (default ``"cuda"``), ``phase`` (``"reference"``:
    the reference's global phase convention; ``"raw"``: the plain projector
    contraction), ``renormalize`` (default True, as the reference normalizes
    every step).  ``force0=False`` draws every measurement outcome and applies
    the flow byproduct corrections (the PennyLane backend's semantics,
    `pennylane_simulator.py:138-153`); ``seed`` seeds the draws and
    ``outcomes`` holds the last run's outcomes per node.
    """

    _exclude = "output"
    _can_sample = True

    def _backend_kwargs(self, kwargs):
        self.phase = kwargs.pop("phase", "reference")
        if self.phase not in ("reference", "raw"):
            raise ValueError("phase must be 'reference' or 'raw'")
        self.renormalize = kwargs.pop("renormalize", True)

    def _check_planes(self, mbqcircuit):
        for node in mbqcircuit.graph.nodes:
            m = mbqcircuit[node]
            if m is not None and m.plane not in ("X", "Y", "XY"):
                raise ValueError(f"Node {node} has plane {m.plane}, but only XY plane is supported.")
        if mbqcircuit.classical_output_nodes:
            raise NotImplementedError("measured output nodes are not supported by hip-sv")

    def run(self, angles, output_form: str = "dm", **kwargs):
        self._check_run(angles, kwargs.get("input_state"))
        form = output_form.lower()
        if form not in ("dm", "densitymatrix", "sv", "statevector"):
            raise ValueError(f"Output form {output_form} is not supported.")
        a, s = self._device_inputs(angles)
        if self.force0:
            out = sv_forward(self.plan, a, s, ref_phase=self.phase == "reference",
                             renormalize=self.renormalize)
            bits = None
        else:
            out, rows = sv_sample(self.plan, a, s, self._next_seed())
            self.last_outcomes = rows
            bits = decode_outcomes(rows[0, 0], self.plan.info["n_steps"])
        psi = out[0, 0].cpu().numpy().astype(np.complex128)
        self._finish_run(bits)
        self.qstate = psi
        if form in ("dm", "densitymatrix"):
            return np.outer(psi, np.conj(psi))
        return psi

    def _window(self, k):
        torch = _torch()
        a, s, W = self._window_args(k)
        out = torch.empty((1, 1, 1 << W), dtype=torch.complex64, device=s.device)
        flags = nat.FLAG_REF_PHASE if self.phase == "reference" else 0
        with torch.cuda.device(s.device):
            nat.check(nat.lib().mbqc_sv_window_c64(
                self.plan.handle, C.c_void_p(a.data_ptr()), 0, C.c_void_p(s.data_ptr()), 0, k,
                C.c_void_p(out.data_ptr()), 1, 1, flags, _stream_ptr(s.device)))
        return out[0, 0].cpu().numpy().astype(np.complex128), 0

    def run_batch(self, angles, input_states=None, *, phase: Optional[str] = None,
                  renormalize: Optional[bool] = None, want_phase: bool = False):
        """All-pairs batch: ``angles [B, n]`` x ``input_states [S, 2^|I|]`` ->
        device tensor ``[B, S, 2^|O|]`` (torch complex64, stays in HBM).
        With ``force0=False`` the outcomes are sampled (``last_outcomes`` holds
        the ``[B, S, W + 2]`` outcome words and output frame)."""
        dev = self.device
        _require_gpu(dev)
        a = _as_angles(angles, self.plan.info["n_angles"], dev)
        st = self._input_np if input_states is None else input_states
        s = _as_states(st, self.plan.info["n_inputs"], dev)
        if not self.force0:
            if want_phase:
                raise ValueError("want_phase is not available with sampled outcomes")
            out, rows = sv_sample(self.plan, a, s, self._next_seed())
            self.last_outcomes = rows
            return out
        ph = self.phase if phase is None else phase
        rn = self.renormalize if renormalize is None else renormalize
        return sv_forward(self.plan, a, s, ref_phase=ph == "reference", renormalize=rn,
                          want_phase=want_phase)


class HipSimulatorDM(_HipBackend):
    """Density-matrix backend with optional noise.

    Extra keywords: ``device``, ``noise`` (a ``mentpy_amd.noise.NoiseModel``),
    or ``p_depol`` / ``p_readout`` floats, or the PennyLane backend's
    ``circuit_noise`` name with ``p`` / ``gamma`` (pennylane_simulator.py:40-46,
    123-136); ``renormalize`` (default True: the reference divides by the
    outcome probability each step).
    """

    _exclude = "quantum_output"
    _plane_error = NotImplementedError
    _all_planes = True

    def _backend_kwargs(self, kwargs):
        from ..noise import NoiseModel
        noise = kwargs.pop("noise", None)
        pd = kwargs.pop("p_depol", None)
        pr = kwargs.pop("p_readout", None)
        cn = kwargs.pop("circuit_noise", None)
        cp = kwargs.pop("p", None)
        cg = kwargs.pop("gamma", None)
        given = sum(x is not None for x in (noise, cn)) + (pd is not None)
        if given > 1:
            raise ValueError("pass one of noise=, p_depol= or circuit_noise=")
        if noise is None:
            if cn is not None:
                noise = NoiseModel.from_circuit_noise(cn, p=cp or 0.0, gamma=cg or 0.0,
                                                      readout_flip=pr or 0.0)
            else:
                noise = NoiseModel(depolarizing=pd or 0.0, readout_flip=pr or 0.0)
        elif pr is not None:
            raise ValueError("pass either noise= or p_depol/p_readout, not both")
        self.noise = noise
        self.renormalize = kwargs.pop("renormalize", True)

    @property
    def _has_z(self):
        return any(self.mbqcircuit[v].plane == "Z" for v in self.schedule_measure)

    @property
    def _exact_probs(self):
        # Every outcome probability is 1/2 on a pattern with a causal flow and
        # XY measurements; otherwise the reference's per-step probabilities
        # (and its outcome-1 rule) are evaluated on the device.
        return (bool(self.plan.info["general_planes"]) or self.mbqcircuit.flow is None
                or not self.plan.info["rotating"])

    def _check_mode(self, mode):
        if self._has_z and mode not in ("expectation", "exp"):
            raise NotImplementedError(
                "Z-plane nodes: in mode 'sample' the reference draws their outcome with an unseeded "
                "np.random.choice (np_simulator_dm.py:323-327); run them with mode='expectation', "
                "which traces them out unprojected")

    def _check_step_mode(self, ment, mode):
        if ment.plane == "Z" and mode not in ("expectation", "exp"):
            self._check_mode(mode)

    def run(self, angles, mode: str = "sample", input_state=None):
        self._check_run(angles, input_state)
        self._check_mode(mode)
        a, s = self._device_inputs(angles)
        if self._exact_probs:
            torch = _torch()
            oc = torch.empty((1, 1, self.plan.info["n_steps"]), dtype=torch.float32, device=s.device)
            rho = dm_forward(self.plan, a, s, noise=self.noise, renormalize=self.renormalize,
                             outcomes=oc)
            vals = oc[0, 0].cpu().numpy()
            self._finish_run()
            self.outcomes = {v: (float(o) if self.mbqcircuit[v].plane == "Z" else int(o))
                             for v, o in zip(self.measure_order, vals)}
        else:
            rho = dm_forward(self.plan, a, s, noise=self.noise, renormalize=self.renormalize)
            self._finish_run()
        out = rho[0, 0].cpu().numpy().astype(np.complex128)
        self.qstate = out
        return out

    def _window(self, k):
        torch = _torch()
        a, s, W = self._window_args(k)
        out = torch.empty((1, 1, 1 << W, 1 << W), dtype=torch.complex64, device=s.device)
        noise = self.noise.native()
        oc = None
        if self._exact_probs and k > 0:
            oc = torch.empty((1, 1, k), dtype=torch.float32, device=s.device)
        with torch.cuda.device(s.device):
            nat.check(nat.lib().mbqc_dm_window_c64(
                self.plan.handle, C.c_void_p(a.data_ptr()), 0, C.c_void_p(s.data_ptr()), 0, k,
                C.c_void_p(out.data_ptr()), 1, 1, C.byref(noise),
                C.c_void_p(oc.data_ptr() if oc is not None else 0), _stream_ptr(s.device)))
        outcome = 0
        if oc is not None:
            last = float(oc[0, 0, k - 1].item())
            node = self.measure_order[k - 1]
            outcome = last if self.mbqcircuit[node].plane == "Z" else int(last)
        return out[0, 0].cpu().numpy().astype(np.complex128), outcome

    def run_batch(self, angles, input_states=None, renormalize: Optional[bool] = None,
                  mode: str = "sample", outcomes=None):
        """``angles [B, n]`` x pure ``input_states [S, 2^|I|]`` -> ``[B, S, d, d]``.
        ``outcomes``: optional float32 ``[B, S, n_steps]`` device tensor that
        receives every simulation's outcomes."""
        self._check_mode(mode)
        dev = self.device
        _require_gpu(dev)
        a = _as_angles(angles, self.plan.info["n_angles"], dev)
        st = self._input_np if input_states is None else input_states
        s = _as_states(st, self.plan.info["n_inputs"], dev)
        rn = self.renormalize if renormalize is None else renormalize
        return dm_forward(self.plan, a, s, noise=self.noise, renormalize=rn,
                          step_probs=self._exact_probs, outcomes=outcomes)
