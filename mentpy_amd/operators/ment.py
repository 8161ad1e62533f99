"""Measurement operators for measurement patterns.

Mirrors the measurement object of mentpy (`mentpy/operators/ment.py:123-263`):
the same constructor forms, plane rules, trainability rule and matrix/POVM
definitions, so code written against mentpy's ``Ment`` runs unchanged.
"""
from __future__ import annotations

import math
import warnings
from typing import Optional, Union

import numpy as np

from .gates import PauliX, PauliY, PauliZ

__all__ = ["Ment", "Measurement"]

_PLANES = ("XY", "XZ", "YZ", "XYZ", "X", "Y", "Z")


class Ment:
    """A single-qubit measurement in a plane of the Bloch sphere.

    ``Ment(angle=None, plane="XY")``; ``Ment("X")`` and ``Ment(plane="X")`` are
    Pauli measurements with stored angle 0.  A measurement whose angle is
    ``None`` in a continuous plane is trainable
    (`mentpy/operators/ment.py:135-173, 214-216`).
    """

    def __init__(self, angle: Optional[Union[int, float, tuple, str]] = None,
                 plane: Optional[str] = "XY"):
        if isinstance(angle, (int, float, tuple)) or angle is None:
            plane = plane if plane is not None else "XY"
        elif isinstance(angle, str):
            swapped_plane = angle
            angle = plane if isinstance(plane, (int, float, tuple)) else None
            plane = swapped_plane
        else:
            raise TypeError(
                f"Invalid argument type. Expected float or str but got {type(angle)}")
        plane = plane.upper()
        if plane not in _PLANES:
            raise ValueError(f"Plane {plane} is not supported.")
        if plane == "XYZ":
            warnings.warn("Plane XYZ might be unstable. Use at your own risk.")
        if plane in ("X", "Y", "Z"):
            if angle is not None and angle != 0:
                raise ValueError(f"Plane {plane} does not support angle.")
            angle = 0
        self._plane = plane
        self._angle = angle
        self._node_id = -1

    def __repr__(self):
        if isinstance(self._angle, tuple):
            theta = tuple(round(a, 4) for a in self._angle)
        elif isinstance(self._angle, (int, float)):
            theta = round(self._angle, 4)
        else:
            theta = "θ"
        return f"Ment({theta}, {self._plane})"

    @property
    def plane(self) -> str:
        return self._plane

    @property
    def angle(self):
        return self._angle

    @property
    def node_id(self):
        return self._node_id

    @node_id.setter
    def node_id(self, node_id):
        self._node_id = node_id

    def set_angle(self, angle):
        self._angle = angle
        return self

    def copy(self) -> "Ment":
        return Ment(self._angle, self._plane)

    def is_trainable(self) -> bool:
        return self._angle is None and self._plane in ("XY", "XZ", "YZ", "XYZ")

    def xy_angle(self, angle: Optional[float] = None) -> float:
        """Equivalent XY-plane angle of this measurement.

        Plane ``X`` is the XY measurement at 0 and plane ``Y`` the XY
        measurement at pi/2 (`mentpy/operators/ment.py:231-235`); a fixed XY
        angle is used as stored.  Other planes raise ``ValueError`` as the
        statevector backend does (`np_simulator_sv.py:54-59`).
        """
        if self._plane == "X":
            return 0.0
        if self._plane == "Y":
            return math.pi / 2
        if self._plane != "XY":
            raise ValueError(f"Plane {self._plane} is not an XY-plane measurement.")
        if self._angle is None:
            if angle is None:
                raise ValueError("Measurement is trainable, please provide an angle.")
            return float(angle)
        return float(self._angle)

    def matrix(self, angle: Optional[float] = None, *args, **kwargs):
        """Observable of the measurement (`mentpy/operators/ment.py:218-253`)."""
        if self._angle is None and angle is None:
            raise ValueError("Measurement is trainable, please provide an angle.")
        if self._angle is not None and angle is not None and self._angle != angle:
            raise ValueError(f"Measurement has a fixed angle of {round(self._angle, 4)}")
        if self._angle is not None:
            angle = self._angle
        if self._plane == "XY":
            return np.cos(angle) * PauliX + np.sin(angle) * PauliY
        if self._plane in ("X", "Y", "Z"):
            return {"X": PauliX, "Y": PauliY, "Z": PauliZ}[self._plane]
        if self._plane == "XZ":
            return np.cos(angle) * PauliX + np.sin(angle) * PauliZ
        if self._plane == "YZ":
            return np.cos(angle) * PauliY + np.sin(angle) * PauliZ
        if not isinstance(angle, tuple):
            raise TypeError(f"Invalid argument type. Expected tuple but got {type(angle)}")
        a1, a2 = angle
        return (np.cos(a1) * np.cos(a2) * PauliX + np.sin(a1) * np.cos(a2) * PauliY
                + np.sin(a2) * PauliZ)

    def get_povm(self, angle: Optional[float] = None, *args, **kwargs):
        """Projectors ``(I +- M)/2`` (`mentpy/operators/ment.py:255-260`)."""
        mat = self.matrix(angle, *args, **kwargs)
        return (np.eye(2) + mat) / 2, (np.eye(2) - mat) / 2


Measurement = Ment
