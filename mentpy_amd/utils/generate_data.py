"""Seeded Haar-random data for MB-QML (`mentpy/utils/generate_data.py:15-57`).

mentpy draws Haar states as the first column of ``random_su`` on the unseeded
global RNG; here every function takes a ``numpy.random.Generator`` (or seed)
so data sets are reproducible.
"""
from __future__ import annotations

import numpy as np

from ..operators.gates import random_su

__all__ = ["generate_haar_random_states", "generate_random_dataset", "train_test_split",
           "random_special_unitary"]


def _rng(rng):
    return rng if isinstance(rng, np.random.Generator) else np.random.default_rng(rng)


def random_special_unitary(n_qubits: int, rng=None) -> np.ndarray:
    return random_su(n_qubits, _rng(rng))


def generate_haar_random_states(n_qubits: int, n_samples: int = 1, rng=None):
    """One state (n_samples == 1) or an ``[n_samples, 2^n]`` array of Haar states."""
    g = _rng(rng)
    states = np.stack([random_su(n_qubits, g)[:, 0] for _ in range(n_samples)])
    return states[0] if n_samples == 1 else states


def train_test_split(inputs, targets, test_size: float = 0.3, randomize: bool = False, rng=None):
    n = len(inputs)
    n_test = int(n * test_size)
    n_train = n - n_test
    inputs, targets = np.asarray(inputs), np.asarray(targets)
    if randomize:
        perm = _rng(rng).permutation(n)
        inputs, targets = inputs[perm], targets[perm]
    return (inputs[:n_train], targets[:n_train]), (inputs[n_train:], targets[n_train:])


def generate_random_dataset(unitary: np.ndarray, n_samples: int, test_size: float = 0.3, rng=None):
    """Haar inputs and targets ``U psi`` split into train and test sets."""
    n_qubits = int(np.log2(unitary.shape[0]))
    inputs = np.atleast_2d(generate_haar_random_states(n_qubits, n_samples, rng))
    targets = inputs @ np.asarray(unitary).T
    return train_test_split(inputs, targets, test_size=test_size)
