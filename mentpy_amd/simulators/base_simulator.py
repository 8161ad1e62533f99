"""Backend interface of the pattern simulator (`mentpy/simulators/base_simulator.py:13-101`)."""
from __future__ import annotations

import abc
from typing import List

__all__ = ["BaseSimulator"]


class BaseSimulator(abc.ABC):
    """A backend holds a pattern, an input state and the outcomes of the last run."""

    def __init__(self, mbqcircuit, input_state=None) -> None:
        self._mbqcirc = mbqcircuit
        self._input_state = input_state
        self._outcomes = {}

    @property
    def mbqcircuit(self):
        return self._mbqcirc

    @property
    def input_state(self):
        return self._input_state

    @input_state.setter
    def input_state(self, input_state):
        self._input_state = input_state

    @property
    def outcomes(self) -> dict:
        return self._outcomes

    @outcomes.setter
    def outcomes(self, outcomes: dict):
        self._outcomes = outcomes

    def __call__(self, angles: List[float], **kwargs):
        return self.run(angles, **kwargs)

    def __repr__(self) -> str:
        return f"{self.__class__.__name__} for {self.mbqcircuit}"

    @abc.abstractmethod
    def measure(self, angle: float, **kwargs):
        """Perform the next measurement of the schedule."""

    @abc.abstractmethod
    def run(self, angles: List[float], **kwargs):
        """Measure the whole pattern with the given angles."""

    @abc.abstractmethod
    def reset(self, input_state=None):
        """Return to the initial state (optionally with a new input state)."""
