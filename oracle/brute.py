"""Full-register simulation of small patterns in gate order (complex128).

Independent of any window logic: the whole resource state is built first
(input on the input nodes, |+> elsewhere, every CZ of the graph), then each
measured node is projected with sqrt(2) <+_theta| in measurement order, as the
PennyLane backend orders the circuit (`mentpy/simulators/pennylane_simulator.py:116-153`,
with outcomes post-selected to 0 so no byproduct correction fires).  The
output is in ``output_nodes`` order, in the raw frame (no global phase
convention): the reference's convention multiplies it by
``prod_k sgn(cos(theta_k/2)) exp(i theta_k/2)`` (see ``closed_forms``).

``brute_dm`` does the same on density matrices with depolarizing noise on
every qubit after all CZs and readout flips at the measurements.  It projects
each node onto the +1 eigenvector of its ``Ment.matrix`` (any plane but XYZ,
`mentpy/operators/ment.py:218-253`, from a numerical eigendecomposition, so it
shares no closed form with the window oracles), takes outcome 1 when outcome
0 has probability below 1e-4 and, with ``mode="expectation"``, traces Z-plane
nodes out unprojected, as `mentpy/simulators/np_simulator_dm.py:307-346` does.
Register size is 2^N (4^N for DM), so only patterns of up to ~14 (7) nodes.
"""
from __future__ import annotations

import numpy as np

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)


def _node_angle(circ, node, angles):
    trainable = list(circ.trainable_nodes)
    if node in trainable:
        return float(angles[trainable.index(node)])
    m = circ[node]
    if m.plane == "X":
        return 0.0
    if m.plane == "Y":
        return np.pi / 2
    return float(m.angle)


def brute_sv(circ, angles, input_state, measured=None):
    """Raw output state (unit norm) of the post-selected pattern."""
    nodes = list(circ.graph.nodes())
    full = brute_graph_state(circ, input_state)
    if measured is None:
        measured = [v for v in circ.measurement_order if v not in circ.output_nodes]
    live = list(nodes)
    for v in measured:
        th = _node_angle(circ, v, angles)
        ax = live.index(v)
        full = np.tensordot(np.array([1.0, np.exp(-1j * th)]), full, axes=([0], [ax]))
        live.pop(ax)
        full /= np.linalg.norm(full)
    out_perm = [live.index(v) for v in circ.output_nodes]
    return np.transpose(full, out_perm).reshape(-1)


def brute_sv_sampled(circ, angles, input_state, outcomes):
    """The PennyLane backend's circuit with given outcomes, literally
    (`mentpy/simulators/pennylane_simulator.py:138-153`): node v is rotated by
    RZ(-theta), H, and projected on |m_v>; an outcome 1 applies X to flow(v)
    and Z to every neighbour of flow(v) that comes after v in the measurement
    order (output nodes included).  ``outcomes`` maps node -> bit.

    Returns ``(state, probability, frame_x, frame_z)``: the normalised output
    state in ``output_nodes`` order, the probability of the outcome string, and
    the X / Z corrections that landed on the output qubits as bit masks
    (bit n_out-1-m for output m), counted classically from the outcomes.
    """
    nodes = list(circ.graph.nodes())
    full = brute_graph_state(circ, input_state)
    order = list(circ.measurement_order)
    measured = [v for v in order if v not in circ.output_nodes]
    live = list(nodes)
    prob = 1.0
    outs = list(circ.output_nodes)
    fx = fz = 0
    for v in measured:
        th = _node_angle(circ, v, angles)
        m = int(outcomes[v])
        ax = live.index(v)
        bra = np.array([1.0, (-1) ** m * np.exp(-1j * th)]) / np.sqrt(2)
        full = np.tensordot(bra, full, axes=([0], [ax]))
        live.pop(ax)
        p = float(np.linalg.norm(full)) ** 2
        prob *= p
        full /= np.sqrt(p)
        if m:
            f = circ.flow(v)
            ka = live.index(f)
            full = np.flip(full, axis=ka).copy()                     # X on flow(v)
            if f in outs:
                fx ^= 1 << (len(outs) - 1 - outs.index(f))
            for nb in circ.graph.neighbors(f):
                if nb != v and order.index(nb) > order.index(v):
                    kb = live.index(nb)
                    idx = [slice(None)] * full.ndim
                    idx[kb] = 1
                    full[tuple(idx)] *= -1                          # Z on nb
                    if nb in outs:
                        fz ^= 1 << (len(outs) - 1 - outs.index(nb))
    out_perm = [live.index(v) for v in outs]
    return np.transpose(full, out_perm).reshape(-1), prob, fx, fz


def _ment_eigvecs(circ, node, angles):
    """(+1, -1) eigenvectors of the node's measurement observable."""
    m = circ[node]
    trainable = list(circ.trainable_nodes)
    th = float(angles[trainable.index(node)]) if node in trainable else m.angle
    if m.plane == "XY":
        M = np.cos(th) * X + np.sin(th) * Y
    elif m.plane in ("X", "Y", "Z"):
        M = {"X": X, "Y": Y, "Z": Z}[m.plane]
    elif m.plane == "XZ":
        M = np.cos(th) * X + np.sin(th) * Z
    elif m.plane == "YZ":
        M = np.cos(th) * Y + np.sin(th) * Z
    else:
        raise NotImplementedError(f"plane {m.plane}")
    w, V = np.linalg.eigh(M)
    return V[:, int(np.argmax(w))], V[:, int(np.argmin(w))]


def brute_dm(circ, angles, input_state, p_depol=0.0, p_readout=0.0, kraus=None, mode="sample"):
    """Output density matrix (quantum_output_nodes order) with noise: the
    depolarizing channel ``p_depol`` and then the Kraus channel ``kraus`` (a list of
    2x2 operators) on every qubit after all CZs, readout flips at the measurements."""
    nodes = list(circ.graph.nodes())
    N = len(nodes)
    psi = brute_graph_state(circ, input_state)
    rho = np.multiply.outer(psi, psi.conj())          # axes: N row, N col
    for k in range(N):
        rho = _depol(rho, k, N, p_depol)
        if kraus is not None:
            rho = _kraus(rho, k, N, kraus)
    measured = [v for v in circ.measurement_order if v not in circ.quantum_output_nodes]
    live = list(nodes)
    for v in measured:
        ax = live.index(v)
        n = len(live)
        tr_all = np.trace(rho.reshape(2 ** n, 2 ** n)).real
        if circ[v].plane == "Z" and mode in ("expectation", "exp"):
            rho = np.trace(rho, axis1=ax, axis2=n + ax) / tr_all
            live.pop(ax)
            continue
        if circ[v].plane == "Z":
            raise NotImplementedError("the reference draws Z-plane outcomes unseeded")
        plus, minus = _ment_eigvecs(circ, v, angles)

        def proj(vec):
            r = np.tensordot(vec.conj(), rho, axes=([0], [ax]))          # row
            return np.tensordot(r, vec, axes=([n - 1 + ax], [0]))        # col (shifted)
        a, b = proj(plus), proj(minus)
        if np.trace(a.reshape(2 ** (n - 1), 2 ** (n - 1))).real / tr_all < 1e-4:
            a, b = b, a                                                  # outcome 1
        mix = (1 - p_readout) * a + p_readout * b
        tr = np.trace(mix.reshape(2 ** (n - 1), 2 ** (n - 1)))
        rho = mix / tr
        live.pop(ax)
    n = len(live)
    outs = list(circ.quantum_output_nodes)
    perm = [live.index(v) for v in outs]
    rho = np.transpose(rho, perm + [n + p for p in perm])
    return rho.reshape(2 ** n, 2 ** n)


def brute_graph_state(circ, input_state):
    nodes = list(circ.graph.nodes())
    N = len(nodes)
    axis = {v: k for k, v in enumerate(nodes)}
    order = list(circ.input_nodes) + [v for v in nodes if v not in circ.input_nodes]
    full = np.asarray(input_state, dtype=complex).reshape([2] * len(circ.input_nodes))
    plus = np.array([1, 1], dtype=complex) / np.sqrt(2)
    for _ in order[len(circ.input_nodes):]:
        full = np.multiply.outer(full, plus)
    full = np.transpose(full, [order.index(v) for v in nodes])
    for a, b in circ.graph.edges():
        idx = [slice(None)] * N
        idx[axis[a]] = 1
        idx[axis[b]] = 1
        full[tuple(idx)] *= -1
    return full


def _kraus(rho, k, N, kraus):
    out = np.zeros_like(rho)
    for K in kraus:
        t = np.moveaxis(np.tensordot(K, rho, axes=([1], [k])), 0, k)
        t = np.moveaxis(np.tensordot(t, K.conj(), axes=([N + k], [1])), -1, N + k)
        out = out + t
    return out


def _depol(rho, k, N, p):
    if p == 0:
        return rho
    out = (1 - p) * rho
    for P in (X, Y, Z):
        t = np.tensordot(P, rho, axes=([1], [k]))
        t = np.moveaxis(t, 0, k)
        t = np.tensordot(t, P.conj().T, axes=([N + k], [0]))
        t = np.moveaxis(t, -1, N + k)
        out = out + (p / 3) * t
    return out
