"""Noise model of the density-matrix path.

mentpy's ``mentpy/noise`` package is an empty stub (`mentpy/noise/base_noise.py:8-12`)
and its numpy backends are noiseless; the only noise in mentpy is the
PennyLane backend's ``circuit_noise``, one channel on every qubit after all CZs
(`mentpy/simulators/pennylane_simulator.py:118-136`).  The model here keeps
that placement for the depolarizing channel
``rho -> (1-p) rho + p/3 (X rho X + Y rho Y + Z rho Z)`` and adds a classical
readout flip of every measured qubit with probability ``readout_flip``
(post-selection is on the reported outcome 0).
"""
from __future__ import annotations

from dataclasses import dataclass

__all__ = ["NoiseModel", "coherence_factor"]


@dataclass(frozen=True)
class NoiseModel:
    depolarizing: float = 0.0
    readout_flip: float = 0.0

    def __post_init__(self):
        for name in ("depolarizing", "readout_flip"):
            v = getattr(self, name)
            if not (0.0 <= v <= 1.0):
                raise ValueError(f"{name} probability must lie in [0, 1], got {v}")


def coherence_factor(noise: NoiseModel) -> float:
    """Damping of the measured qubit's XY coherence: (1 - 4p/3)(1 - 2 p_m).

    Depolarizing shrinks the Bloch vector by (1 - 4p/3); a readout flip of the
    post-selected outcome mixes the two XY projectors, shrinking it by
    (1 - 2 p_m).
    """
    return (1.0 - 4.0 * noise.depolarizing / 3.0) * (1.0 - 2.0 * noise.readout_flip)
