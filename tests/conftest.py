import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def haar_state(n_qubits, rng):
    """Haar-random pure state: the first column of a Haar unitary, or (above
    10 qubits, where the unitary's QR would dominate a test) a normalised
    complex Gaussian vector, which has the same distribution."""
    if n_qubits > 10:
        v = rng.normal(size=2 ** n_qubits) + 1j * rng.normal(size=2 ** n_qubits)
        return v / np.linalg.norm(v)
    from mentpy_amd.operators.gates import haar_unitary
    return haar_unitary(2 ** n_qubits, rng)[:, 0]


def phase_aligned_max_err(a, b):
    """max |a - e^{i phi} b| after aligning the global phase (per last axis row)."""
    a = np.asarray(a)
    b = np.asarray(b)
    ov = np.sum(np.conj(b) * a, axis=-1, keepdims=True)
    ph = ov / np.maximum(np.abs(ov), 1e-300)
    return float(np.max(np.abs(a - ph * b)))


def infidelity(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    na = np.sum(np.abs(a) ** 2, -1)
    nb = np.sum(np.abs(b) ** 2, -1)
    return 1.0 - np.abs(np.sum(np.conj(a) * b, -1)) ** 2 / (na * nb)
