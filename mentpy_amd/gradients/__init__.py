"""Gradients of pattern costs: mentpy's black-box methods and the GPU adjoint."""
from .adjoint import (AdjointEngine, MBQMLStep, PatternSVFunction, fd_loop_infidelity_grad, mbqml_infidelity,
                      pattern_sv)
from .grad import AdjointCost, fd_gradient, get_gradient, get_hessian, psr_gradient

__all__ = ["get_gradient", "get_hessian", "fd_gradient", "psr_gradient", "AdjointCost",
           "AdjointEngine", "MBQMLStep", "PatternSVFunction", "pattern_sv", "mbqml_infidelity",
           "fd_loop_infidelity_grad"]
