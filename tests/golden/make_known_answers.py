"""Write known_answers.json: the literal input/expected pairs of mentpy's own
known-answer tests for the simulate path, as data.

Sources (mentpy 0.1.0a9):
  tests/test_calculator.py:12-54   pure2density and partial_trace examples
  tests/operators/test_ment.py:31-65  Ment matrices of the XY / X / Y / Z planes
  tests/test_simulators.py:15-30   teleportation: a linear cluster of 2i+1
                                   nodes measured at angle 0 returns the input
The values are the ones those tests state; complex arrays are stored as
[real, imag] pairs.  Run from the repository root:
  python tests/golden/make_known_answers.py
"""
import json
from pathlib import Path

import numpy as np


def c(a):
    a = np.asarray(a, dtype=complex)
    return {"shape": list(a.shape), "re": a.real.ravel().tolist(), "im": a.imag.ravel().tolist()}


X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
s = 1 / np.sqrt(2)
plus = np.array([s, s])
rho_mixed = np.array([[0.5, 0], [0, 0.5]])
sigma = np.outer(plus, plus)

cases = {
    "pure2density": [
        {"psi": c([1, 0]), "rho": c([[1, 0], [0, 0]])},
        {"psi": c([s, s]), "rho": c([[0.5, 0.5], [0.5, 0.5]])},
    ],
    "partial_trace": [
        {"state": c(rho_mixed), "indices": [0], "expected": c(1.0)},
        {"state": c(np.kron(rho_mixed, sigma)), "indices": [0], "expected": c(sigma)},
        {"state": c(np.kron(rho_mixed, sigma)), "indices": [1], "expected": c(rho_mixed)},
        {"state": c([s, s]), "indices": [0], "expected": c(1.0)},
        {"state": c(np.kron([s, s], plus)), "indices": [0], "expected": c(plus)},
        {"state": c(np.kron([s, s], plus)), "indices": [1], "expected": c([s, s])},
    ],
    "ment_matrix": [
        {"plane": "XY", "angle": float(np.pi / 4),
         "matrix": c(np.cos(np.pi / 4) * X + np.sin(np.pi / 4) * Y)},
        {"plane": "X", "angle": 0.0, "matrix": c(X)},
        {"plane": "Y", "angle": 0.0, "matrix": c(Y)},
        {"plane": "Z", "angle": 0.0, "matrix": c(Z)},
    ],
    "teleportation": [{"n_nodes": 2 * i + 1, "angles": [0.0] * (2 * i), "atol": 1e-3}
                      for i in range(1, 5)],
}

out = Path(__file__).with_name("known_answers.json")
out.write_text(json.dumps(cases, indent=1))
print(f"wrote {out}")
