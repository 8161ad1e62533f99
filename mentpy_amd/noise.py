"""Noise model of the density-matrix path.

mentpy's ``mentpy/noise`` package is an empty stub (`mentpy/noise/base_noise.py:8-12`)
and its numpy backends are noiseless; the only noise in mentpy is the
PennyLane backend's ``circuit_noise``: one channel on every qubit after all CZs
(`mentpy/simulators/pennylane_simulator.py:118-136`), one of
``"depolarizing"``, ``"amplitude_damping"``, ``"phase_damping"``,
``"phase_flip"``, ``"generalized_amplitude_damping"``.  ``NoiseModel`` keeps
that placement for each of these channels (several may be combined; they act
in the field order below) and adds a classical readout flip of every measured
qubit with probability ``readout_flip`` (post-selection is on the reported
outcome 0).

Channels (Kraus forms as documented for the PennyLane operations):
``depolarizing`` p: rho -> (1-p) rho + p/3 (X rho X + Y rho Y + Z rho Z);
``amplitude_damping`` gamma; ``phase_damping`` gamma; ``phase_flip`` p;
``gad_gamma``, ``gad_p``: generalized amplitude damping (gamma, p).
"""
from __future__ import annotations

from dataclasses import dataclass, fields

__all__ = ["NoiseModel", "coherence_factor", "CIRCUIT_NOISE"]

CIRCUIT_NOISE = ("depolarizing", "amplitude_damping", "phase_damping", "phase_flip",
                 "generalized_amplitude_damping")


@dataclass(frozen=True)
class NoiseModel:
    depolarizing: float = 0.0
    readout_flip: float = 0.0
    amplitude_damping: float = 0.0
    phase_damping: float = 0.0
    phase_flip: float = 0.0
    gad_gamma: float = 0.0
    gad_p: float = 0.0

    def __post_init__(self):
        for f in fields(self):
            v = getattr(self, f.name)
            if not (0.0 <= v <= 1.0):
                raise ValueError(f"{f.name} must lie in [0, 1], got {v}")

    @classmethod
    def from_circuit_noise(cls, circuit_noise, p=0.0, gamma=0.0, readout_flip=0.0):
        """The PennyLane backend's ``circuit_noise`` names (pennylane_simulator.py:123-136):
        ``p`` is the probability of depolarizing / phase flip (and the
        excitation probability of generalized amplitude damping), ``gamma`` the
        damping rate of the damping channels."""
        if circuit_noise is None:
            return cls(readout_flip=readout_flip)
        if circuit_noise == "depolarizing":
            return cls(depolarizing=p, readout_flip=readout_flip)
        if circuit_noise == "amplitude_damping":
            return cls(amplitude_damping=gamma, readout_flip=readout_flip)
        if circuit_noise == "phase_damping":
            return cls(phase_damping=gamma, readout_flip=readout_flip)
        if circuit_noise == "phase_flip":
            return cls(phase_flip=p, readout_flip=readout_flip)
        if circuit_noise == "generalized_amplitude_damping":
            return cls(gad_gamma=gamma, gad_p=p, readout_flip=readout_flip)
        raise ValueError(f"Unrecognized circuit noise: {circuit_noise}")

    def native(self):
        from . import _native as nat
        return nat.Noise(p_depol=self.depolarizing, p_readout=self.readout_flip,
                         amp_damping=self.amplitude_damping, phase_damping=self.phase_damping,
                         phase_flip=self.phase_flip, gad_gamma=self.gad_gamma, gad_p=self.gad_p)


def coherence_factor(noise: NoiseModel) -> float:
    """Damping of a measured qubit's XY coherence by the noise and the readout
    flip: the product of each channel's off-diagonal factor (depolarizing
    1 - 4p/3, damping channels sqrt(1 - gamma), phase flip 1 - 2p) and 1 - 2 p_m.
    """
    import math
    c = (1.0 - 4.0 * noise.depolarizing / 3.0)
    c *= math.sqrt(1.0 - noise.amplitude_damping) * math.sqrt(1.0 - noise.phase_damping)
    c *= (1.0 - 2.0 * noise.phase_flip) * math.sqrt(1.0 - noise.gad_gamma)
    return c * (1.0 - 2.0 * noise.readout_flip)
