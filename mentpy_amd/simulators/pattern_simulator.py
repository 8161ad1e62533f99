"""Backend-dispatching facade (`mentpy/simulators/pattern_simulator.py:19-88`).

The reference dispatches over a closed dict of backends; here the dict is open
(``register_backend``) and ships the MI355X backends ``"hip-sv"`` and
``"hip-dm"``.  Other backends (for example a CPU reference used by tests) can
be registered under their own names.  The default input state is
``|+>^{|I|}``; torch tensors are accepted as input states besides numpy arrays.
"""
from __future__ import annotations

from typing import Dict, List, Type

import numpy as np

from .base_simulator import BaseSimulator

__all__ = ["PatternSimulator", "register_backend", "available_backends"]

_BACKENDS: Dict[str, Type[BaseSimulator]] = {}


def register_backend(name: str, cls: Type[BaseSimulator]) -> None:
    """Make ``cls`` selectable as ``PatternSimulator(..., backend=name)``."""
    if not (isinstance(cls, type) and issubclass(cls, BaseSimulator)):
        raise TypeError(f"{cls} is not a BaseSimulator subclass")
    _BACKENDS[name.lower()] = cls


def available_backends() -> List[str]:
    return sorted(_BACKENDS)


def _is_state(x) -> bool:
    if isinstance(x, np.ndarray):
        return True
    try:
        import torch
        return isinstance(x, torch.Tensor)
    except ImportError:  # pragma: no cover
        return False


class PatternSimulator:
    """Simulator for measurement patterns.

    ``PatternSimulator(mbqcircuit, input_state=None, backend="hip-sv", *args, **kwargs)``
    instantiates the backend with ``(mbqcircuit, input_state, *args, **kwargs)``.
    """

    def __init__(self, mbqcircuit, input_state=None, backend: str = "hip-sv", *args, **kwargs):
        backend = backend.lower()
        if backend not in _BACKENDS:
            raise ValueError(f"Backend {backend} not supported. Supported backends are "
                             f"{_BACKENDS.keys()}")
        if input_state is None:
            input_state = np.ones(1)
            for _ in range(len(mbqcircuit.input_nodes)):
                input_state = np.kron(input_state, np.array([1, 1]) / np.sqrt(2))
        elif not _is_state(input_state):
            raise TypeError(f"Input state must be a numpy array, not {type(input_state)}")
        self.simulator = _BACKENDS[backend](mbqcircuit, input_state, *args, **kwargs)

    def __getattr__(self, name):
        if name == "simulator":
            raise AttributeError(name)
        return getattr(self.simulator, name)

    def __call__(self, angles, **kwargs):
        return self.run(angles, **kwargs)

    def __repr__(self) -> str:
        return f"{self.__class__.__name__} ({self.simulator!r})"

    def measure(self, angle: float, **kwargs):
        return self.simulator.measure(angle, **kwargs)

    def run(self, angles, **kwargs):
        return self.simulator.run(angles, **kwargs)

    def reset(self, input_state=None):
        return self.simulator.reset(input_state)
