"""Resource-graph templates used by the simulate path.

``linear_cluster``, ``many_wires``, ``grid_cluster``, ``muta`` and ``spturb``
build the same graphs, node numbering, edge insertion order and I/O nodes as
mentpy (`mentpy/mbqc/templates.py:16-294`); insertion order matters because it
fixes ``trainable_nodes`` (the angle layout) and the tie order of the schedule.
"""
from __future__ import annotations

from typing import List

from ..operators.ment import Ment
from .graphstate import GraphState
from .mbqcircuit import MBQCircuit, hstack, merge

__all__ = ["linear_cluster", "many_wires", "grid_cluster", "muta", "spturb"]


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


def muta(n_wires: int, n_layers: int, **kwargs) -> MBQCircuit:
    """Multiple Triangle Ansatz (`mentpy/mbqc/templates.py:150-216`): per wire a
    block of ``n_wires`` 5-node wires whose second node is joined to the first
    and third node of every other wire; blocks are stacked horizontally over
    wires, then over layers.  As in mentpy, each ``add_edge`` recompiles the
    pattern, which resets the measurements (so ``restrict_trainable`` has no
    lasting effect)."""
    options = {"restrict_trainable": True, "one_column": False}
    options.update(kwargs)
    size = 5
    big = None
    for wire in range(n_wires):
        if options["one_column"] and wire != 0:
            break
        g = many_wires([size] * n_wires)
        if options["restrict_trainable"]:
            g.trainable_nodes = list(set(g.trainable_nodes) - set(i - 1 for i in g.output_nodes))
        for connect in range(n_wires):
            if connect != wire:
                g.add_edge(size * wire + 1, size * connect)
                g.add_edge(size * wire + 1, size * connect + 2)
        big = g if big is None else hstack((big, g))
    out = None
    for _ in range(n_layers):
        out = big if out is None else hstack((out, big))
    return out


def spturb(n_qubits: int, n_layers: int, periodic: bool = False, **kwargs) -> MBQCircuit:
    """Symmetry Protected Topological Perturbator Ansatz
    (`mentpy/mbqc/templates.py:219-294`): 3-node wires with a trainable middle
    node, and per layer two rounds of 12-node symmetry blocks merged onto
    output pairs (i, i+2)."""
    if n_qubits < 4:
        raise ValueError("n_qubits must be greater than 4")
    gr = many_wires([5, 2, 5]).graph
    gr.add_edge(2, 6)
    gr.add_edge(9, 6)
    block1 = MBQCircuit(gr, input_nodes=[0, 7], output_nodes=[4, 11],
                        measurements={5: Ment(plane="XY")}, default_measurement=Ment(plane="X"))
    block2 = MBQCircuit(gr, input_nodes=[0, 7], output_nodes=[4, 11],
                        measurements={5: Ment(plane="XY"), 1: Ment(plane="Y"), 3: Ment(plane="Y"),
                                      8: Ment(plane="Y"), 10: Ment(plane="Y")},
                        default_measurement=Ment(plane="X"))

    def layer():
        return many_wires([3] * n_qubits, measurements={3 * i + 1: Ment() for i in range(n_qubits)},
                          default_measurement=Ment(plane="X"))
    spt = layer()
    n_blocks = n_qubits if periodic else n_qubits - 2
    for lay in range(n_layers):
        if lay != 0:
            spt = hstack((spt, layer()))
        for m in range(2):
            for i in range(n_blocks):
                n1 = spt.output_nodes[i]
                n2 = spt.output_nodes[(i + 2) % n_qubits]
                spt = merge(spt, block1 if m % 2 == 0 else block2, [(n1, 0), (n2, 7)])
    return spt
