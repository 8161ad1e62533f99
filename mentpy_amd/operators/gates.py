"""Single-qubit gate constants and Haar-random unitaries.

Only the pieces of mentpy's gate module that sit on the simulate-with-angles
path are provided: the Pauli/Hadamard constants used by ``Ment`` and the Haar
sampler used to make input states (`mentpy/operators/gates.py:16-34`).  The
dense kron-chain operator builders of mentpy are not needed: the HIP kernels
apply CZ signs and projections directly on amplitudes.
"""
from __future__ import annotations

import numpy as np

__all__ = ["HGate", "PauliX", "PauliY", "PauliZ", "SGate", "TGate", "random_su"]

HGate = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
PauliX = np.array([[0, 1], [1, 0]])
PauliY = np.array([[0, -1j], [1j, 0]])
PauliZ = np.array([[1, 0], [0, -1]])
SGate = np.array([[1, 0], [0, 1j]])
TGate = np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]])


def haar_unitary(dim: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-random ``dim x dim`` unitary from a seeded generator.

    QR of a complex Ginibre matrix with the diagonal phase fix (Mezzadri,
    "How to generate random matrices from the classical compact groups",
    Notices AMS 54 (2007)), which is the distribution
    ``scipy.stats.unitary_group`` samples.
    """
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))


def random_su(n_qubits: int, rng: np.random.Generator | None = None) -> np.ndarray:
    """Haar-random special unitary on ``n_qubits`` (`mentpy/operators/gates.py:29-34`).

    The reference divides a Haar unitary by ``det(U)**(1/2**n)``; the same is
    done here, with an explicit generator so results are reproducible.
    """
    rng = np.random.default_rng() if rng is None else rng
    dim = 2 ** n_qubits
    u = haar_unitary(dim, rng)
    return u / np.power(np.linalg.det(u), 1.0 / dim)
