"""Resource-graph templates used by the simulate path.

``linear_cluster``, ``many_wires`` and ``grid_cluster`` build the same graphs,
node numbering, edge insertion order and I/O nodes as mentpy
(`mentpy/mbqc/templates.py:16-147`); insertion order matters because it fixes
``trainable_nodes`` (the angle layout) and the tie order of the schedule.
"""
from __future__ import annotations

from typing import List

from .graphstate import GraphState
from .mbqcircuit import MBQCircuit

__all__ = ["linear_cluster", "many_wires", "grid_cluster"]


def linear_cluster(n: int, **kwargs) -> MBQCircuit:
    """1D cluster of ``n`` qubits, input node 0, output node ``n-1``."""
    g = GraphState()
    g.add_edges_from([(i, i + 1) for i in range(n - 1)])
    return MBQCircuit(g, input_nodes=[0], output_nodes=[n - 1], **kwargs)


def many_wires(n_wires: List[int], **kwargs) -> MBQCircuit:
    """Disjoint 1D wires of the given lengths (each at least 2)."""
    if not all(isinstance(n, int) and n > 1 for n in n_wires):
        raise ValueError("n_wires must be a list of integers greater than 1")
    g = GraphState()
    start = 0
    for n in n_wires:
        g.add_edges_from([(start + j, start + j + 1) for j in range(n - 1)])
        start += n
    starts = [sum(n_wires[:i]) for i in range(len(n_wires))]
    ends = [sum(n_wires[:i + 1]) - 1 for i in range(len(n_wires))]
    return MBQCircuit(g, input_nodes=starts, output_nodes=ends, **kwargs)


def grid_cluster(n: int, m: int, periodic: bool = False, **kwargs) -> MBQCircuit:
    """``n`` rows by ``m`` columns cluster; node ``i*m + j`` is row i, column j.

    Row edges are inserted first (row by row), then column edges, so graph
    order is node-id order.  Inputs are column 0, outputs column ``m-1``.
    """
    g = GraphState()
    for i in range(n):
        g.add_edges_from([(i * m + j, i * m + j + 1) for j in range(m - 1)])
    for i in range(n - 1):
        g.add_edges_from([(i * m + j, (i + 1) * m + j) for j in range(m)])
    if periodic and n > 1:
        for j in range(m):
            g.add_edge(j, (n - 1) * m + j)
    return MBQCircuit(g, input_nodes=[i * m for i in range(n)],
                      output_nodes=[i * m + m - 1 for i in range(n)], **kwargs)
