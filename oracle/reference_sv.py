"""Step-by-step restatement of mentpy's ``NumpySimulatorSV`` (complex128).

Follows `mentpy/simulators/np_simulator_sv.py` line by line in structure:

* ``__init__`` (:38-128): window = |I|+1 when left at 1 and a flow exists,
  input reordered from ``input_nodes`` to ``schedule[:|I|]`` order, padded with
  |+>, multiplied by the dense product of CZs inside the first window.
* ``measure`` (:164-225): project window qubit 0 (the MSB), renormalise,
  amplitude-sum partial trace (`mentpy/calculator/state_ops.py:42-74`), append
  |+> as the new LSB and apply dense CZs to its window neighbours.
* ``measure_ment`` (:322-358): POVM ``(I +- M)/2`` with ``M = cos X + sin Y``
  (`mentpy/operators/ment.py:218-260`) embedded by a kron chain
  (`mentpy/operators/gates.py:62-72`); outcome forced to 0.
* ``run`` (:227-297): angle of node i is ``angles[trainable_nodes.index(i)]``,
  fixed-angle nodes use their own angle; output reordered to ``output_nodes``;
  returns ``outer(psi, psi*)`` or ``psi``.
* ``dev_mode`` (:130-162, 176-197, 254-276): the window is a list of nodes,
  initially ``schedule[:w]``; each step measures the first listed node whose
  first later neighbour on its wire (``wires``) is listed too (or that has
  none), at its list index, drops it from the list and appends the next
  scheduled node as before.

Dense operators are built exactly as the reference builds them (kron chains of
2x2 factors), so the cost profile of this class is the reference's; it is the
timed CPU baseline.  Pinned by ``tests/test_oracle.py`` (teleportation and
random-parity known-answer tests of mentpy's own suite, chain closed form,
brute force on small grids).
"""
from __future__ import annotations

from typing import List

import numpy as np

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
I2 = np.eye(2)
PLUS = np.array([1, 1]) / np.sqrt(2)


def embed(u, i, n):
    """``u`` on qubit ``i`` of ``n`` (MSB first) -- kron chain (gates.py:62-72)."""
    op = np.ones((1, 1))
    for k in range(n):
        op = np.kron(op, u if k == i else I2)
    return op


def controlled_z(i, j, n):
    """Dense CZ built as ``I - 2 |11><11|`` by kron chains (gates.py:127-143)."""
    one = np.array([[0, 0], [0, 1]])
    op1, op2 = np.ones((1, 1)), 2 * np.ones((1, 1))
    for k in range(n):
        op1 = np.kron(op1, I2)
        op2 = np.kron(op2, one if k in (i, j) else I2)
    return op1 - op2


def swap_ij(i, j, n):
    """Dense SWAP of qubits i and j (gates.py:75-98)."""
    d = 2 ** n
    op = np.zeros((d, d))
    for idx in range(d):
        bi = (idx >> (n - 1 - i)) & 1
        bj = (idx >> (n - 1 - j)) & 1
        out = idx & ~(1 << (n - 1 - i)) & ~(1 << (n - 1 - j))
        out |= bj << (n - 1 - i)
        out |= bi << (n - 1 - j)
        op[out, idx] = 1
    return op


def find_swaps(source, target):
    """Transposition sequence taking ``source`` order to ``target`` (np_simulator_sv.py:360-374)."""
    assert set(source) == set(target), "Both lists must have the same elements"
    swaps, src = [], list(source)
    for i, t in enumerate(target):
        if src[i] != t:
            j = src.index(t, i + 1)
            src[i], src[j] = src[j], src[i]
            swaps.append((i, j))
    return swaps


def reorder_qubits(state, current, target):
    """Apply the swap chain (np_simulator_sv.py:376-384 / np_simulator_dm.py:364-380)."""
    new = np.array(state, dtype=complex, copy=True)
    for i, j in find_swaps(current, target):
        sw = swap_ij(i, j, len(current))
        new = sw @ new if new.ndim == 1 else sw @ new @ sw.conj().T
    return new


def partial_trace_pure_state(psi, indices):
    """Amplitude-sum "partial trace" of a pure state, then normalise (state_ops.py:42-74)."""
    n = int(np.log2(psi.shape[0]))
    remaining = sorted(set(range(n)) - set(indices))
    t = psi.reshape([2] * n).transpose(remaining + list(indices))
    t = t.reshape([-1] + [2] * len(indices)).sum(axis=tuple(range(1, 1 + len(indices))))
    t = t.reshape(-1)
    return t / np.linalg.norm(t)


def partial_trace_mixed_state(rho, indices):
    """Sum over basis isometries of the traced qubits (state_ops.py:77-119)."""
    zero, one = np.array([[1], [0]]), np.array([[0], [1]])
    n = int(np.log2(rho.shape[0]))
    r = len(indices)
    sigma = np.zeros((2 ** (n - r), 2 ** (n - r)), dtype=complex)
    for m in range(2 ** r):
        bits = format(m, f"0{r}b")
        iso = np.ones((1, 1))
        for k in range(n):
            if k in indices:
                iso = np.kron(iso, zero if bits[indices.index(k)] == "0" else one)
            else:
                iso = np.kron(iso, I2)
        sigma += iso.conj().T @ rho @ iso
    return sigma


def pure2density(psi):
    return np.outer(psi, np.conj(psi))


def partial_trace(data, indices):
    if data.ndim == 1:
        return partial_trace_pure_state(data, indices)
    if data.ndim == 2:
        return partial_trace_mixed_state(data, indices)
    raise ValueError("Invalid input shape for quantum state.")


def ment_matrix(plane, stored_angle, angle):
    """Observable of a measurement (ment.py:218-253) for the planes on this path."""
    if stored_angle is not None:
        angle = stored_angle
    if plane == "XY":
        return np.cos(angle) * X + np.sin(angle) * Y
    if plane in ("X", "Y", "Z"):
        return {"X": X, "Y": Y, "Z": Z}[plane]
    if plane == "XZ":
        return np.cos(angle) * X + np.sin(angle) * Z
    if plane == "YZ":
        return np.cos(angle) * Y + np.sin(angle) * Z
    raise ValueError(f"plane {plane} not supported by the oracle")


def povm(plane, stored_angle, angle):
    m = ment_matrix(plane, stored_angle, angle)
    return (np.eye(2) + m) / 2, (np.eye(2) - m) / 2


def _node_angle(circ, node, angles):
    trainable = list(circ.trainable_nodes)
    if node in trainable:
        return angles[trainable.index(node)]
    return circ[node].angle


class ReferenceSV:
    """Faithful complex128 restatement of ``NumpySimulatorSV`` (force0 only)."""

    def __init__(self, circ, input_state=None, window_size=1, schedule=None, dev_mode=False,
                 wires=None):
        self.circ = circ
        self.window_size = window_size
        self.schedule = schedule
        self.dev_mode = dev_mode
        self.wires = wires
        for node in circ.graph.nodes:
            m = circ[node]
            if m is not None and m.plane not in ("X", "Y", "XY"):
                raise ValueError(f"Node {node} has plane {m.plane}, but only XY plane is supported.")
        if self.schedule is not None:
            self.schedule_measure = [i for i in self.schedule if i not in circ.output_nodes]
        elif circ.measurement_order is not None:
            self.schedule_measure = [i for i in circ.measurement_order if i not in circ.output_nodes]
            self.schedule = list(circ.measurement_order)
            if self.window_size == 1 and circ.flow is not None:
                self.window_size = len(circ.input_nodes) + 1
        else:
            raise ValueError("Schedule must be provided.")
        if input_state is None:
            input_state = np.ones(1)
            for _ in circ.input_nodes:
                input_state = np.kron(input_state, PLUS)
        n_in = len(circ.input_nodes)
        if n_in > self.window_size:
            raise ValueError("Input state must have at most as many qubits as the window size.")
        if self.window_size > len(self.schedule_measure):
            raise ValueError("Window size larger than the number of measurements.")
        self.n_nodes = len(circ.graph.nodes)
        w = self.window_size
        sub = circ.graph.subgraph(self.schedule[:w])
        self.initial_czs = np.eye(2 ** w)
        window = self.schedule[:w]
        for a in sub.nodes:
            for b in sub.neighbors(a):
                if a < b:
                    self.initial_czs = controlled_z(window.index(a), window.index(b), w) @ self.initial_czs
        self.reset(input_state)

    def current_simulated_nodes(self) -> List:
        if self.dev_mode:
            return self._current_simulated_nodes
        return self.schedule[self.current_measurement:self.current_measurement + self.window_size]

    def _wire_of(self, node):
        for wire in self.wires:
            if node in wire:
                return wire
        raise TypeError(f"node {node} is on none of the wires")

    def _dev_pick(self):
        """First window node whose first later wire neighbour is in the window
        (or that has none): (node, index), or (None, None)."""
        window = self.current_simulated_nodes()
        for node in window:
            wire = self._wire_of(node)
            later = [n for n in wire if self.circ.graph.has_edge(node, n) and n > node]
            if not later or later[0] in window:
                return node, window.index(node)
        return None, None

    def current_number_simulated_nodes(self) -> int:
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
        self.qstate = self.initial_czs @ self.input_state
        self.outcomes = {}

    def measure_ment(self, plane, stored, angle, i):
        p0, p1 = povm(plane, stored, angle)
        n = self.current_number_simulated_nodes()
        p0e, p1e = embed(p0, i, n), embed(p1, i, n)
        self.prob0 = np.vdot(self.qstate, p0e @ self.qstate)
        self.prob1 = np.vdot(self.qstate, p1e @ self.qstate)
        st = p0e @ self.qstate
        return st / np.linalg.norm(st)

    def measure(self, angle):
        if self.current_measurement >= len(self.schedule_measure):
            raise ValueError("No more measurements to be done.")
        if self.dev_mode:
            node, indx = self._dev_pick()
            if node is None:
                raise ValueError("no window node can be measured")
        else:
            node, indx = self.schedule_measure[self.current_measurement], 0
        m = self.circ[node]
        self.qstate = self.measure_ment(m.plane, m.angle, angle, indx)
        self.current_measurement += 1
        self.qstate = partial_trace(self.qstate, [indx])
        if self.dev_mode:
            self._current_simulated_nodes = [v for v in self._current_simulated_nodes if v != node]
            if self.current_measurement + self.window_size <= self.n_nodes:
                self._current_simulated_nodes.append(
                    self.schedule[self.current_measurement + self.window_size - 1])
        if self.current_measurement + self.window_size <= self.n_nodes:
            self.qstate = np.kron(self.qstate, PLUS)
            window = self.current_simulated_nodes()
            new = window[-1]
            i_new = window.index(new)
            for nbr in self.circ.graph.neighbors(new):
                if nbr in window:
                    self.qstate = controlled_z(window.index(nbr), i_new, self.window_size) @ self.qstate
        return self.qstate, 0

    def run(self, angles, output_form="dm", input_state=None):
        if input_state is not None:
            self.reset(input_state)
        if len(angles) != len(self.circ.trainable_nodes):
            raise ValueError("Number of angles does not match number of trainable nodes.")
        if self.dev_mode:
            while self.current_measurement < len(self.schedule_measure):
                node, _ = self._dev_pick()
                if node is None:
                    raise ValueError("no window node can be measured")
                self.qstate, out = self.measure(_node_angle(self.circ, node, angles))
                self.outcomes[node] = out
        else:
            for node in self.schedule_measure:
                self.qstate, out = self.measure(_node_angle(self.circ, node, angles))
                self.outcomes[node] = out
        current = self.current_simulated_nodes()
        if list(self.circ.quantum_output_nodes) != current:
            self.qstate = reorder_qubits(self.qstate, current, self.circ.output_nodes)
        if output_form.lower() in ("dm", "densitymatrix"):
            return np.outer(self.qstate, np.conj(self.qstate))
        if output_form.lower() in ("sv", "statevector"):
            return self.qstate
        raise ValueError(f"Output form {output_form} is not supported.")
