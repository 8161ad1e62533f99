"""Gate constants, dense n-qubit operator builders and Haar-random unitaries.

The names and conventions of mentpy's gate module
(`mentpy/operators/gates.py:10-143`): qubit 0 is the most significant bit (the
leftmost kron factor); ``cnot_ij`` numbers qubits from 1, as mentpy's does,
while ``arbitrary_qubit_gate``, ``swap_ij`` and ``controlled_z`` number them
from 0.  The HIP kernels never build these matrices (they apply CZ signs and
projections to amplitudes); the builders are here for code written against
mentpy that uses them.  They are built from index arithmetic (diagonals and
permutations) rather than kron chains.
"""
from __future__ import annotations

import numpy as np

__all__ = ["CNOT", "SWAP", "CSGate", "HGate", "PauliX", "PauliY", "PauliZ", "SGate", "TGate",
           "random_su", "arbitrary_qubit_gate", "swap_ij", "cnot_ij", "controlled_z"]

CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])
CSGate = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1j]])

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


def _bit(idx, q, n):
    """Value of qubit q (0 = most significant) in basis indices idx of n qubits."""
    return (idx >> (n - 1 - q)) & 1


def arbitrary_qubit_gate(u, i, n):
    """The one-qubit gate ``u`` on qubit ``i`` of ``n`` (identity elsewhere),
    `mentpy/operators/gates.py:62-72`."""
    u = np.asarray(u)
    return np.kron(np.kron(np.eye(2 ** i), u), np.eye(2 ** (n - i - 1)))


def swap_ij(i, j, n):
    """Permutation exchanging qubits ``i`` and ``j`` of ``n``,
    `mentpy/operators/gates.py:75-98`."""
    assert i < n and j < n
    idx = np.arange(2 ** n)
    bi, bj = _bit(idx, i, n), _bit(idx, j, n)
    src = idx ^ ((bi ^ bj) << (n - 1 - i)) ^ ((bi ^ bj) << (n - 1 - j))
    op = np.zeros((2 ** n, 2 ** n))
    op[src, idx] = 1.0
    return op


def cnot_ij(i, j, n):
    """CNOT with control ``i`` and target ``j`` of ``n`` qubits numbered from
    1, `mentpy/operators/gates.py:101-124`."""
    idx = np.arange(2 ** n)
    flip = _bit(idx, i - 1, n) << (n - j)
    op = np.zeros((2 ** n, 2 ** n))
    op[idx ^ flip, idx] = 1.0
    return op


def controlled_z(i, j, n):
    """CZ between qubits ``i`` and ``j`` of ``n`` (diagonal, -1 where both are
    1), `mentpy/operators/gates.py:127-143`."""
    assert i < n and j < n
    idx = np.arange(2 ** n)
    return np.diag(1.0 - 2.0 * (_bit(idx, i, n) & _bit(idx, j, n)))
