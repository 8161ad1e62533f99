"""Step-by-step restatement of mentpy's ``NumpySimulatorDM`` plus the noise model.

Structure follows `mentpy/simulators/np_simulator_dm.py`:

* ``__init__`` (:33-115): window, input reorder, |+> padding, purification
  ``pure2density`` (`state_ops.py:16-39`), dense initial CZ conjugation.
* ``measure`` (:151-216): ``measure_ment`` on window qubit 0, NaN guard,
  ``partial_trace_mixed_state`` (`state_ops.py:77-119`), append |+><+|,
  dense CZ conjugation with window neighbours.
* ``measure_ment`` (:307-346): ``prob = Re tr(rho P_ext)``, outcome 0 unless
  ``prob0 < 1e-4`` (then 1), ``rho <- P rho P^dag / prob``; any plane of
  ``Ment.matrix`` but XYZ (`mentpy/operators/ment.py:218-253`).  A Z-plane
  node in ``mode="expectation"`` is not projected: its outcome is
  ``prob1 / (prob0 + prob1)`` and the qubit is traced out (:318-322, 339-340).
  In ``mode="sample"`` the reference draws a Z outcome with an unseeded
  ``np.random.choice`` (:323-327); here a draw needs an explicit ``rng``.
* ``run`` (:218-283): schedule without ``quantum_output_nodes``, output
  reordered to ``quantum_output_nodes``.
* ``dev_mode`` (:151-216, 240-262): the window is a node list; each step
  measures the first listed node whose first later wire neighbour is listed
  (or that has none), at its list index (as in ``reference_sv``).

Noise (absent from mentpy's numpy backends; placement of the PennyLane backend,
`pennylane_simulator.py:118-136`): every qubit is depolarized with ``p`` after
all of its CZs.  In window form this is just before the qubit is measured (the
channel commutes with everything applied to other qubits in between), and at
the end for output qubits.  A readout flip ``p_m`` makes the post-selected
update ``((1-p_m) P0 rho P0 + p_m P1 rho P1) / ((1-p_m) prob0 + p_m prob1)``.
Dense Kraus/projector operators are kron-built as in the reference.

``kraus`` (a list of 2x2 Kraus operators, ``oracle.channels``) adds any other
single-qubit channel at the same place, after the depolarizing one.

Pinned by ``tests/test_oracle.py``: p = 0 equals the SV oracle's density
matrix; with noise it equals the brute-force Kraus simulation in gate order
(``oracle.brute.brute_dm``) on small grids; teleportation known-answer test.
"""
from __future__ import annotations

import numpy as np

from .reference_sv import (PLUS, X, Y, Z, _node_angle, controlled_z, embed,
                           partial_trace_mixed_state, povm, pure2density, reorder_qubits)


def depolarize(rho, i, n, p):
    """(1-p) rho + p/3 sum_P P_i rho P_i on qubit i of n."""
    if p == 0:
        return rho
    out = (1 - p) * rho
    for P in (X, Y, Z):
        Pe = embed(P, i, n)
        out = out + (p / 3) * (Pe @ rho @ Pe.conj().T)
    return out


def apply_channel(rho, i, n, kraus):
    """sum_K K_i rho K_i^dag on qubit i of n (dense kron-embedded operators)."""
    if kraus is None:
        return rho
    out = np.zeros_like(rho)
    for K in kraus:
        Ke = embed(K, i, n)
        out = out + Ke @ rho @ Ke.conj().T
    return out


class ReferenceDM:
    """Faithful complex128 restatement of ``NumpySimulatorDM`` (force0 only)
    with optional depolarizing / readout-flip noise."""

    def __init__(self, circ, input_state=None, window_size=1, schedule=None,
                 p_depol=0.0, p_readout=0.0, kraus=None, dev_mode=False, wires=None):
        self.circ = circ
        self.dev_mode = dev_mode
        self.wires = wires
        self.kraus = kraus
        self.window_size = window_size
        self.schedule = schedule
        self.p_depol = float(p_depol)
        self.p_readout = float(p_readout)
        qout = list(circ.quantum_output_nodes)
        if self.schedule is not None:
            self.schedule_measure = [i for i in self.schedule if i not in qout]
        elif circ.measurement_order is not None:
            self.schedule_measure = [i for i in circ.measurement_order if i not in qout]
            self.schedule = list(circ.measurement_order)
            if self.window_size == 1 and circ.flow is not None:
                self.window_size = len(circ.input_nodes) + 1
        else:
            raise ValueError("Schedule must be provided.")
        if input_state is None:
            input_state = np.ones(1)
            for _ in circ.input_nodes:
                input_state = np.kron(input_state, PLUS)
        if len(circ.input_nodes) > self.window_size:
            raise ValueError("Input state must have at most as many qubits as the window size.")
        if self.window_size > len(self.schedule_measure):
            raise ValueError("Window size larger than the number of measurements.")
        self.n_nodes = len(circ.graph.nodes)
        w = self.window_size
        window = self.schedule[:w]
        sub = circ.graph.subgraph(window)
        self.initial_czs = np.eye(2 ** w)
        for a in sub.nodes:
            for b in sub.neighbors(a):
                if a < b:
                    self.initial_czs = controlled_z(window.index(a), window.index(b), w) @ self.initial_czs
        self.reset(input_state)

    def current_simulated_nodes(self):
        if self.dev_mode:
            return self._current_simulated_nodes
        return self.schedule[self.current_measurement:self.current_measurement + self.window_size]

    def _dev_pick(self):
        """dev_mode (np_simulator_dm.py:151-175, 240-262): the first window
        node whose first later wire neighbour is in the window (or that has
        none), as (node, index); (None, None) if there is none."""
        window = self.current_simulated_nodes()
        for node in window:
            wire = next((w for w in self.wires if node in w), None)
            if wire is None:
                raise TypeError(f"node {node} is on none of the wires")
            later = [n for n in wire if self.circ.graph.has_edge(node, n) and n > node]
            if not later or later[0] in window:
                return node, window.index(node)
        return None, None

    def current_number_simulated_nodes(self):
        return min(self.window_size, self.n_nodes - self.current_measurement)

    def reset(self, input_state=None):
        self.current_measurement = 0
        if self.dev_mode:
            self._current_simulated_nodes = list(self.schedule[:self.window_size])
        if input_state is not None:
            st = reorder_qubits(np.asarray(input_state, dtype=complex), self.circ.input_nodes,
                                self.schedule[:len(self.circ.input_nodes)])
            for _ in range(self.window_size - len(self.circ.input_nodes)):
                st = np.kron(st, PLUS)
            self.input_state = st
        rho = pure2density(self.input_state)
        self.qstate = self.initial_czs @ rho @ self.initial_czs.conj().T
        self.outcomes = {}

    def measure_ment(self, plane, stored, angle, i, mode="sample", rng=None):
        n = self.current_number_simulated_nodes()
        rho = depolarize(self.qstate, i, n, self.p_depol)
        rho = apply_channel(rho, i, n, self.kraus)
        p0, p1 = povm(plane, stored, angle)
        p0e, p1e = embed(p0, i, n), embed(p1, i, n)
        prob0 = np.real(np.trace(rho @ p0e))
        prob1 = np.real(np.trace(rho @ p1e))
        if plane == "Z":
            if mode in ("expectation", "exp"):
                return rho, prob1 / (prob0 + prob1)
            if rng is None:
                raise NotImplementedError("the reference draws Z-plane outcomes unseeded; pass rng")
            outcome = int(rng.choice([0, 1], p=np.array([prob0, prob1]) / (prob0 + prob1)))
        else:
            outcome = 1 if prob0 < 1e-4 else 0
        pa, pb = (p0e, p1e) if outcome == 0 else (p1e, p0e)
        qa, qb = (prob0, prob1) if outcome == 0 else (prob1, prob0)
        pm = self.p_readout
        num = (1 - pm) * (pa @ rho @ pa.conj().T) + pm * (pb @ rho @ pb.conj().T)
        return num / ((1 - pm) * qa + pm * qb), outcome

    def measure(self, angle, mode="sample", rng=None):
        if self.current_measurement >= len(self.schedule_measure):
            raise ValueError("No more measurements to be done.")
        if self.dev_mode:
            node, indx = self._dev_pick()
            if node is None:
                raise ValueError("no window node can be measured")
        else:
            node, indx = self.schedule_measure[self.current_measurement], 0
        m = self.circ[node]
        self.qstate, outcome = self.measure_ment(m.plane, m.angle, angle, indx, mode, rng)
        if np.isnan(self.qstate).any():
            raise ValueError("qstate has nan, you might want to increase the window size")
        self.current_measurement += 1
        self.qstate = partial_trace_mixed_state(self.qstate, [indx])
        if self.dev_mode:
            self._current_simulated_nodes = [v for v in self._current_simulated_nodes if v != node]
            if self.current_measurement + self.window_size <= self.n_nodes:
                self._current_simulated_nodes.append(
                    self.schedule[self.current_measurement + self.window_size - 1])
        if self.current_measurement + self.window_size <= self.n_nodes:
            self.qstate = np.kron(self.qstate, pure2density(PLUS))
            window = self.current_simulated_nodes()
            new = window[-1]
            i_new = window.index(new)
            for nbr in self.circ.graph.neighbors(new):
                if nbr in window:
                    cz = controlled_z(window.index(nbr), i_new, self.window_size)
                    self.qstate = cz @ self.qstate @ cz.conj().T
        return self.qstate, outcome

    def run(self, angles, mode="sample", input_state=None, rng=None):
        if input_state is not None:
            self.reset(input_state)
        if len(angles) != len(self.circ.trainable_nodes):
            raise ValueError("Number of angles does not match number of trainable nodes.")
        if self.dev_mode:
            while self.current_measurement < len(self.schedule_measure):
                node, _ = self._dev_pick()
                if node is None:
                    raise ValueError("no window node can be measured")
                self.qstate, out = self.measure(_node_angle(self.circ, node, angles), mode, rng)
                self.outcomes[node] = out
        else:
            for node in self.schedule_measure:
                self.qstate, out = self.measure(_node_angle(self.circ, node, angles), mode, rng)
                self.outcomes[node] = out
        current = [i for i in self.schedule if i not in self.schedule_measure]
        n = len(current)
        for i in range(n):
            self.qstate = depolarize(self.qstate, i, n, self.p_depol)
            self.qstate = apply_channel(self.qstate, i, n, self.kraus)
        qout = list(self.circ.quantum_output_nodes)
        if qout != current:
            self.qstate = reorder_qubits(self.qstate, current, qout)
        return self.qstate
