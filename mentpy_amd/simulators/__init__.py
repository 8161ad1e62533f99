"""Pattern simulators: the facade, the backend interface and the HIP backends."""
from .base_simulator import BaseSimulator
from .hip_simulator import HipSimulatorDM, HipSimulatorSV, compile_pattern, dm_forward, sv_forward
from .pattern_simulator import PatternSimulator, available_backends, register_backend

register_backend("hip-sv", HipSimulatorSV)
register_backend("hip-dm", HipSimulatorDM)

__all__ = ["BaseSimulator", "PatternSimulator", "HipSimulatorSV", "HipSimulatorDM",
           "register_backend", "available_backends", "compile_pattern", "sv_forward", "dm_forward"]
