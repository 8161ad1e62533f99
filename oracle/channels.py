"""Single-qubit noise channels as Kraus operators (complex128, test oracle).

The PennyLane backend of mentpy puts one channel on every qubit after all CZs
(`mentpy/simulators/pennylane_simulator.py:123-136`): DepolarizingChannel,
AmplitudeDamping, PhaseDamping, PhaseFlip, GeneralizedAmplitudeDamping.
PennyLane is not installed, so the channels are written from their textbook
Kraus forms (Nielsen & Chuang, "Quantum Computation and Quantum Information",
sections 8.3.3-8.3.6; the same forms PennyLane documents for these
operations).  ``tests/test_oracle.py`` pins them by completeness
(sum K^dag K = I), fixed points and closed-form actions.

``noise_kraus`` composes the nonzero channels in the order depolarizing,
amplitude damping, phase damping, phase flip, generalized amplitude damping
(the first listed acts first).
"""
from __future__ import annotations

import numpy as np

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)


def depolarizing(p):
    return [np.sqrt(1 - p) * I2] + [np.sqrt(p / 3) * P for P in (X, Y, Z)]


def amplitude_damping(g):
    return [np.array([[1, 0], [0, np.sqrt(1 - g)]], dtype=complex),
            np.array([[0, np.sqrt(g)], [0, 0]], dtype=complex)]


def phase_damping(g):
    return [np.array([[1, 0], [0, np.sqrt(1 - g)]], dtype=complex),
            np.array([[0, 0], [0, np.sqrt(g)]], dtype=complex)]


def phase_flip(p):
    return [np.sqrt(1 - p) * I2, np.sqrt(p) * Z]


def generalized_amplitude_damping(g, p):
    return [np.sqrt(p) * np.array([[1, 0], [0, np.sqrt(1 - g)]], dtype=complex),
            np.sqrt(p) * np.array([[0, np.sqrt(g)], [0, 0]], dtype=complex),
            np.sqrt(1 - p) * np.array([[np.sqrt(1 - g), 0], [0, 1]], dtype=complex),
            np.sqrt(1 - p) * np.array([[0, 0], [np.sqrt(g), 0]], dtype=complex)]


def compose(first, second):
    """Kraus set of ``second`` applied after ``first``."""
    return [b @ a for a in first for b in second]


def noise_kraus(p_depol=0.0, amp_damping=0.0, phase_damping_=0.0, phase_flip_=0.0,
                gad_gamma=0.0, gad_p=0.0):
    ks = [I2]
    if p_depol:
        ks = compose(ks, depolarizing(p_depol))
    if amp_damping:
        ks = compose(ks, amplitude_damping(amp_damping))
    if phase_damping_:
        ks = compose(ks, phase_damping(phase_damping_))
    if phase_flip_:
        ks = compose(ks, phase_flip(phase_flip_))
    if gad_gamma:
        ks = compose(ks, generalized_amplitude_damping(gad_gamma, gad_p))
    return ks


def is_identity(kraus) -> bool:
    return len(kraus) == 1 and np.allclose(kraus[0], I2, atol=0)


def apply_1q(kraus, rho2):
    """Channel on a single 2x2 block matrix (any linear extension)."""
    return sum(K @ rho2 @ K.conj().T for K in kraus)
