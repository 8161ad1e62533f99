"""mentpy_amd: MI355X-native simulator for measurement-based quantum computing patterns.

Keeps the ``PatternSimulator`` / ``MBQCircuit`` call surface of mentpy and runs
the simulate-with-angles path in hand-written HIP kernels (gfx950) behind a C
ABI (``include/mbqc_hip.h``).
"""
from . import gradients, mbqc, noise, operators, optimizers, utils
from .mbqc import GraphState, MBQCircuit, hstack, templates, vstack
from .mbqc.flow import find_cflow
from .noise import NoiseModel
from .operators import Ment, Measurement, gates
from .simulators import PatternSimulator, register_backend

__version__ = "0.1.0"

__all__ = ["GraphState", "MBQCircuit", "Ment", "Measurement", "PatternSimulator", "NoiseModel",
           "templates", "gates", "find_cflow", "register_backend", "hstack", "vstack"]
