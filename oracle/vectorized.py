"""The reference's per-step operations vectorised over a batch (complex128).

Same window layout and the same operations as ``reference_sv`` /
``reference_dm`` (projector on window qubit 0, normalisation, amplitude-sum
partial trace for SV, |+> append as LSB, CZ with window neighbours), applied
to ``[batch, 2^n]`` arrays with diagonal CZ sign vectors instead of dense
matrices.  Used to check the GPU kernels on thousands of simulations at once.
Pinned in ``tests/test_oracle.py`` against ``reference_sv`` / ``reference_dm``.
"""
from __future__ import annotations

import numpy as np


class _Setup:
    def __init__(self, circ, window_size=1, schedule=None, exclude="output"):
        excl = list(circ.output_nodes if exclude == "output" else circ.quantum_output_nodes)
        if schedule is None:
            schedule = list(circ.measurement_order)
            if window_size == 1 and circ.flow is not None:
                window_size = len(circ.input_nodes) + 1
        self.schedule = list(schedule)
        self.sched_meas = [i for i in self.schedule if i not in excl]
        self.w = window_size
        self.N = len(circ.graph.nodes)
        self.circ = circ
        trainable = list(circ.trainable_nodes)
        self.angle_src = []
        self.planes = []        # "XY" (also X, Y), "XZ", "YZ" or "Z"
        for v in self.sched_meas:
            m = circ[v]
            self.planes.append("XY" if m.plane in ("X", "Y", "XY") else m.plane)
            if v in trainable:
                self.angle_src.append(("t", trainable.index(v)))
            else:
                a = 0.0 if m.plane == "X" else (np.pi / 2 if m.plane == "Y" else float(m.angle))
                self.angle_src.append(("f", a))

    def angles_for_step(self, k, angles):
        kind, val = self.angle_src[k]
        if kind == "t":
            return angles[:, val]
        return np.full(angles.shape[0], val)

    def cz_diag(self, n, pairs):
        idx = np.arange(2 ** n)
        d = np.ones(2 ** n)
        for a, b in pairs:
            d[((idx >> (n - 1 - a)) & 1) & ((idx >> (n - 1 - b)) & 1) == 1] *= -1
        return d

    def initial_pairs(self):
        window = self.schedule[:self.w]
        g = self.circ.graph
        return [(window.index(a), window.index(b)) for a, b in g.subgraph(window).edges()]

    def append_pairs(self, cm):
        window = self.schedule[cm:cm + self.w]
        new = window[-1]
        return [(window.index(nb), window.index(new)) for nb in self.circ.graph.neighbors(new)
                if nb in window]

    def reorder_input(self, states):
        cur = list(self.circ.input_nodes)
        tgt = self.schedule[:len(cur)]
        n = len(cur)
        st = states.reshape([-1] + [2] * n)
        perm = [0] + [1 + cur.index(t) for t in tgt]
        return np.transpose(st, perm).reshape(states.shape[0], -1)


def sv_batch(circ, angles, states, window_size=1, schedule=None, phase="reference"):
    """``angles [B, n]``, ``states [S, 2^|I|]`` -> ``[B, S, 2^|O|]`` (complex128).

    ``phase="reference"`` applies the amplitude-sum partial trace literally
    (mentpy's convention); ``"raw"`` uses the plain ``sqrt(2) <+theta|``
    contraction and divides by the norm.
    """
    su = _Setup(circ, window_size, schedule, "output")
    angles = np.atleast_2d(np.asarray(angles, dtype=np.float64))
    states = np.atleast_2d(np.asarray(states, dtype=complex))
    B, S = angles.shape[0], states.shape[0]
    nin = len(circ.input_nodes)
    st = su.reorder_input(states)
    plus = np.array([1, 1]) / np.sqrt(2)
    for _ in range(su.w - nin):
        st = np.kron(st, plus)
    st = st * su.cz_diag(su.w, su.initial_pairs())
    psi = np.broadcast_to(st[None], (B, S, st.shape[1])).reshape(B * S, -1).copy()
    for k in range(len(su.sched_meas)):
        th = np.repeat(su.angles_for_step(k, angles), S)
        n = int(np.log2(psi.shape[1]))
        p = psi.reshape(B * S, 2, -1)
        e = np.exp(1j * th)[:, None]
        if phase == "reference":
            # P0 = 1/2 [[1, e^{-i th}], [e^{i th}, 1]] on qubit 0, normalise, sum over qubit 0
            a0 = 0.5 * (p[:, 0] + np.conj(e) * p[:, 1])
            a1 = 0.5 * (e * p[:, 0] + p[:, 1])
            nrm = np.sqrt((np.abs(a0) ** 2 + np.abs(a1) ** 2).sum(1))[:, None]
            t = (a0 + a1) / nrm
        else:
            t = p[:, 0] + np.conj(e) * p[:, 1]
        t = t / np.linalg.norm(t, axis=1)[:, None]
        cm = k + 1
        if cm + su.w <= su.N:
            t = np.repeat(t, 2, axis=1) / np.sqrt(2)
            t = t * su.cz_diag(su.w, su.append_pairs(cm))
        psi = t
    cur = su.schedule[len(su.sched_meas):len(su.sched_meas) + su.w]
    tgt = list(circ.output_nodes)
    n = len(cur)
    psi = psi.reshape([B * S] + [2] * n)
    psi = np.transpose(psi, [0] + [1 + cur.index(v) for v in tgt]).reshape(B, S, -1)
    return psi


def _kraus_axis(x, kraus, a, b):
    """sum_K K x K^dag on the (row axis a, column axis b) pair of ``x``."""
    out = np.zeros_like(x)
    for K in kraus:
        t = np.moveaxis(np.tensordot(K, x, axes=([1], [a])), 0, a)
        t = np.moveaxis(np.tensordot(t, K.conj(), axes=([b], [1])), -1, b)
        out = out + t
    return out


def _bloch(plane, th):
    """Bloch vector (n_x, n_y, n_z) of the +1 outcome of ``Ment.matrix``
    (`mentpy/operators/ment.py:218-253`) in a continuous plane; Z is (0, 0, 1)."""
    c, s = np.cos(th), np.sin(th)
    z = np.zeros_like(c)
    if plane == "XY":
        return c, s, z
    if plane == "XZ":
        return c, z, s
    if plane == "YZ":
        return z, c, s
    if plane == "Z":
        return z, z, np.ones_like(c)
    raise NotImplementedError(f"plane {plane}")


def dm_batch(circ, angles, states, window_size=1, schedule=None, p_depol=0.0, p_readout=0.0,
             kraus=None, mode="sample"):
    """``angles [B, n]``, pure ``states [S, 2^|I|]`` -> ``[B, S, d, d]`` with noise
    (depolarizing ``p_depol``, then the Kraus channel ``kraus`` if given, on
    every qubit after its CZs; readout flip ``p_readout``).  Every plane but
    XYZ: the projector ``(I + n.sigma)/2`` of the node's Bloch vector, outcome
    1 where outcome 0 has probability below 1e-4 (per simulation), Z-plane
    nodes traced out unprojected with ``mode="expectation"``
    (`mentpy/simulators/np_simulator_dm.py:307-346`)."""
    su = _Setup(circ, window_size, schedule, "quantum_output")
    angles = np.atleast_2d(np.asarray(angles, dtype=np.float64))
    states = np.atleast_2d(np.asarray(states, dtype=complex))
    B, S = angles.shape[0], states.shape[0]
    nin = len(circ.input_nodes)
    st = su.reorder_input(states)
    plus = np.array([1, 1]) / np.sqrt(2)
    for _ in range(su.w - nin):
        st = np.kron(st, plus)
    st = st * su.cz_diag(su.w, su.initial_pairs())
    psi = np.broadcast_to(st[None], (B, S, st.shape[1])).reshape(B * S, -1)
    rho = psi[:, :, None] * np.conj(psi[:, None, :])
    p, pm = float(p_depol), float(p_readout)
    for k in range(len(su.sched_meas)):
        th = np.repeat(su.angles_for_step(k, angles), S)
        d = rho.shape[1]
        h = d // 2
        r = rho.reshape(B * S, 2, h, 2, h)
        B00, B01, B10, B11 = r[:, 0, :, 0], r[:, 0, :, 1], r[:, 1, :, 0], r[:, 1, :, 1]
        if p:
            # depolarize qubit 0: populations mix by 2p/3, coherences shrink by 1-4p/3
            B00, B11 = (1 - 2 * p / 3) * B00 + (2 * p / 3) * B11, (1 - 2 * p / 3) * B11 + (2 * p / 3) * B00
            B01, B10 = (1 - 4 * p / 3) * B01, (1 - 4 * p / 3) * B10
        if kraus is not None:
            blk = np.stack([np.stack([B00, B01], 1), np.stack([B10, B11], 1)], 1)  # [N,2,2,h,h]
            blk = _kraus_axis(blk, kraus, 1, 2)
            B00, B01, B10, B11 = blk[:, 0, 0], blk[:, 0, 1], blk[:, 1, 0], blk[:, 1, 1]
        plane = su.planes[k]
        if plane == "Z" and mode not in ("expectation", "exp"):
            raise NotImplementedError("the reference draws Z-plane outcomes unseeded")
        if plane == "Z":
            num = B00 + B11
        else:
            nx, ny, nz = (a[:, None, None] for a in _bloch(plane, th))
            e = nx + 1j * ny
            # sum_ab P0[b, a] B_ab with P0 = (I + n.sigma)/2
            plus_part = 0.5 * ((1 + nz) * B00 + (1 - nz) * B11 + e * B01 + np.conj(e) * B10)
            minus_part = 0.5 * ((1 - nz) * B00 + (1 + nz) * B11 - e * B01 - np.conj(e) * B10)
            tr_in = np.real(np.trace(B00 + B11, axis1=1, axis2=2))[:, None, None]
            p0 = np.real(np.trace(plus_part, axis1=1, axis2=2))[:, None, None] / tr_in
            one = p0 < 1e-4
            plus_part, minus_part = (np.where(one, minus_part, plus_part),
                                     np.where(one, plus_part, minus_part))
            num = (1 - pm) * plus_part + pm * minus_part
        tr = np.real(np.trace(num, axis1=1, axis2=2))[:, None, None]
        rho = num / tr
        cm = k + 1
        if cm + su.w <= su.N:
            rho = np.kron(rho, np.full((1, 2, 2), 0.5))
            dg = su.cz_diag(su.w, su.append_pairs(cm))
            rho = rho * dg[None, :, None] * dg[None, None, :]
    cur = [i for i in su.schedule if i not in su.sched_meas]
    n = len(cur)
    dmat = rho.reshape([B * S] + [2] * (2 * n))
    if p:
        for q in range(n):
            dmat = np.moveaxis(dmat, (1 + q, 1 + n + q), (1, 2))
            a, dd = dmat[:, 0, 0].copy(), dmat[:, 1, 1].copy()
            dmat[:, 0, 0] = (1 - 2 * p / 3) * a + (2 * p / 3) * dd
            dmat[:, 1, 1] = (1 - 2 * p / 3) * dd + (2 * p / 3) * a
            dmat[:, 0, 1] *= (1 - 4 * p / 3)
            dmat[:, 1, 0] *= (1 - 4 * p / 3)
            dmat = np.moveaxis(dmat, (1, 2), (1 + q, 1 + n + q))
    if kraus is not None:
        for q in range(n):
            dmat = _kraus_axis(dmat, kraus, 1 + q, 1 + n + q)
    tgt = list(circ.quantum_output_nodes)
    perm = [cur.index(v) for v in tgt]
    dmat = np.transpose(dmat, [0] + [1 + q for q in perm] + [1 + n + q for q in perm])
    return dmat.reshape(B, S, 2 ** n, 2 ** n)
