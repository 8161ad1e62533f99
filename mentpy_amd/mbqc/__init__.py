"""Pattern model: resource graphs, causal flow, measurement patterns, templates."""
from . import templates
from .flow import Flow, check_if_flow, find_cflow
from .graphstate import GraphState
from .mbqcircuit import MBQCircuit, hstack, vstack

__all__ = ["GraphState", "MBQCircuit", "Flow", "find_cflow", "check_if_flow",
           "templates", "hstack", "vstack"]
