"""Causal flow of an open graph and the layering it induces.

The measurement schedule of a pattern comes from its causal flow: the
Mhalla-Perdrix algorithm ("Finding optimal flows efficiently", arXiv:0709.2670)
peels the graph from the outputs backwards, giving every node a layer; nodes
are measured in decreasing layer.  This restates mentpy's ``find_cflow`` /
``causal_flow_aux`` (`mentpy/mbqc/flow.py:115-185`) and the ``Flow`` wrapper
(`mentpy/mbqc/flow.py:58-109`).  Only causal flow is provided: generalised and
Pauli flow (`mentpy/mbqc/flow.py:191-325`) are outside the simulate path, and a
graph without causal flow gets no schedule (the simulator then needs an explicit
``schedule=``), as in mentpy when no flow is found.
"""
from __future__ import annotations

from typing import Callable, Dict, Optional, Tuple

__all__ = ["find_cflow", "find_cflow_native", "Flow", "check_if_flow"]


def find_cflow(graph, input_nodes, output_nodes) -> Tuple[Optional[Callable], Optional[Callable], Optional[int], Optional[Dict]]:
    """Return ``(flow, partial_order, depth, layers)`` or four ``None``.

    ``flow(v)`` is the flow successor of a measured node, ``partial_order(u, v)``
    is ``layer[u] > layer[v]`` (u is measured before v), ``layers`` maps every
    node to its layer (outputs are layer 0).  ``depth`` is the value mentpy
    returns in that slot (the largest successor id, `flow.py:147`).
    """
    if len(input_nodes) != len(output_nodes):
        raise ValueError(
            f"Cannot find flow or gflow. Input ({len(input_nodes)}) and output "
            f"({len(output_nodes)}) nodes have different size.")

    layer = {v: 0 for v in graph.nodes()}
    pending = {v: 0 for v in graph.nodes()}
    succ: Dict = {}
    processed = set(output_nodes)          # grows as nodes are given a successor
    inputs = set(input_nodes)
    all_nodes = set(graph.nodes())

    frontier = set()
    non_out = all_nodes - set(output_nodes)
    for v in set(output_nodes) - inputs:
        pending[v] = len(set(graph.neighbors(v)) & non_out)
        if pending[v] == 1:
            frontier.add(v)

    k = 1
    while True:
        nxt = set()
        for v in frontier:
            cand = set(graph.neighbors(v)) & (all_nodes - processed)
            if len(cand) != 1:
                continue
            u = cand.pop()
            succ[u] = v
            layer[u] = k
            processed.add(u)
            if u not in inputs:
                pending[u] = len(set(graph.neighbors(u)) & (all_nodes - processed))
                if pending[u] == 1:
                    nxt.add(u)
            for x in set(graph.neighbors(u)):
                if pending[x] > 0:
                    pending[x] -= 1
                    if pending[x] == 1:
                        nxt.add(x)
        if not nxt:
            break
        frontier = nxt
        k += 1

    if len(succ) != len(all_nodes) - len(output_nodes):
        return None, None, None, None
    depth = max(succ.values()) if succ else 0
    return (lambda x: succ[x]), (lambda u, v: layer[u] > layer[v]), depth, layer


def find_cflow_native(graph, input_nodes, output_nodes):
    """``find_cflow`` computed by the native library (``mbqc_causal_flow``).

    Returns ``(flow, partial_order, depth, layers, measurement_order)`` with the
    same meaning as ``find_cflow`` plus the ``calculate_order`` schedule, or
    five ``None`` when the graph has no causal flow.  The frontier of each round
    is visited in node order.
    """
    import ctypes as C
    import numpy as np
    from .. import _native as nat
    nodes = list(graph.nodes())
    idx = {v: k for k, v in enumerate(nodes)}
    n = len(nodes)
    edges = np.array([(idx[a], idx[b]) for a, b in graph.edges()], np.int32).reshape(-1, 2)
    ins = np.array([idx[v] for v in input_nodes], np.int32)
    outs = np.array([idx[v] for v in output_nodes], np.int32)
    flow = np.zeros(n, np.int32)
    layer = np.zeros(n, np.int32)
    order = np.zeros(n, np.int32)
    found = C.c_int32(0)
    ptr = lambda a: C.c_void_p(a.ctypes.data)
    nat.check(nat.lib().mbqc_causal_flow(n, len(edges), ptr(edges), len(ins), ptr(ins), len(outs),
                                         ptr(outs), ptr(flow), ptr(layer), ptr(order),
                                         C.byref(found)))
    if not found.value:
        return None, None, None, None, None
    succ = {nodes[u]: nodes[f] for u, f in enumerate(flow) if f >= 0}
    layers = {nodes[u]: int(l) for u, l in enumerate(layer)}
    depth = max(succ.values()) if succ else 0
    return ((lambda x: succ[x]), (lambda u, v: layers[u] > layers[v]), depth, layers,
            [nodes[k] for k in order])


class Flow:
    """Causal flow object: successor function, partial order, layers, order."""

    def __init__(self, graph, input_nodes, output_nodes):
        self.graph = graph
        self.input_nodes = input_nodes
        self.output_nodes = output_nodes
        func, order, depth, layers = find_cflow(graph, input_nodes, output_nodes)
        self.func = func
        self.partial_order = order
        self.depth = depth
        self.layers_dict = layers
        if layers is None:
            self.layers = None
            self.measurement_order = None
            return
        top = max(layers.values())
        grouped = [[n for n, l in layers.items() if l == j] for j in range(top + 1)]
        self.layers = grouped[::-1]
        order_list = [n for group in self.layers for n in group]
        for i in list(self.input_nodes)[::-1]:
            order_list.remove(i)
            order_list.insert(0, i)
        self.measurement_order = order_list

    def __repr__(self):
        return f"Flow(n={self.graph.number_of_nodes()})"

    def __call__(self, node):
        return self.func(node)


def check_if_flow(graph, input_nodes, output_nodes, flow, partial_order) -> bool:
    """Check the three causal-flow conditions (`mentpy/mbqc/flow.py:603-623`)."""
    out = set(output_nodes)
    for i in graph.nodes():
        if i in out:
            continue
        f_i = flow(i)
        if not graph.has_edge(i, f_i) or not partial_order(i, f_i):
            return False
        for k in graph.neighbors(f_i):
            if k != i and not partial_order(i, k):
                return False
    return True
