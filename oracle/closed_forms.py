"""Closed forms that pin the oracle's conventions.

* Chain teleportation: on ``linear_cluster(n)`` every post-selected XY
  measurement at angle theta applies ``J(-theta) = H diag(1, e^{-i theta})`` to the
  logical qubit, so the output is ``H P(-t_m) ... H P(-t_1) psi``; this is the
  gate-model equivalence of the PennyLane backend's ``RZ(-p), H``
  (`mentpy/simulators/pennylane_simulator.py:144-145`).
* Reference phase: mentpy's amplitude-sum partial trace
  (`mentpy/calculator/state_ops.py:68-73`) multiplies the projected state by
  ``(1 + e^{i theta}) / |1 + e^{i theta}| = sgn(cos(theta/2)) e^{i theta/2}`` at
  every step, so reference output = raw output * prod_k of that factor.
"""
from __future__ import annotations

import numpy as np

H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)


def step_phase(theta):
    c = np.cos(theta / 2)
    return np.where(c < 0, -1.0, 1.0) * np.exp(1j * theta / 2)


def reference_phase(thetas):
    out = 1.0 + 0j
    for t in np.atleast_1d(thetas):
        out = out * step_phase(t)
    return out


def chain_output(thetas, psi):
    """Raw output of a post-selected linear cluster (input node 0, output last)."""
    v = np.asarray(psi, dtype=complex)
    for t in thetas:
        v = H @ (np.array([1.0, np.exp(-1j * t)]) * v)
    return v
