"""Restatement of mentpy's schedule derivation for checking the product's.

``cflow_layers`` follows ``find_cflow`` / ``causal_flow_aux``
(`mentpy/mbqc/flow.py:115-185`); ``measurement_order`` follows
``MBQCircuit.calculate_order`` (`mentpy/mbqc/mbqcircuit.py:390-422`) with the
partial order ``layer[u] > layer[v]`` (`flow.py:147`); ``trainable_nodes``
follows the default-measurement rule (`mbqcircuit.py:110-113, 325-353`).
Written as a recursive restatement (like the reference) so it shares no code
with ``mentpy_amd.mbqc``.  Pinned by the grid closed form: layers are columns
(the peeling removes one column per round, outputs first), so the order is
column-major with rows ascending.
"""
from __future__ import annotations

import numpy as np


def cflow_layers(graph, inputs, outputs):
    nodes = list(graph.nodes())
    layer = {v: 0 for v in nodes}
    past = {v: 0 for v in nodes}
    flow = {}
    V = set(nodes)
    outs = set(outputs)
    C = set()
    for v in set(outputs) - set(inputs):
        past[v] = len(set(graph.neighbors(v)) & (V - set(outputs)))
        if past[v] == 1:
            C.add(v)

    def aux(C, k):
        nxt = set()
        for v in C:
            inter = set(graph.neighbors(v)) & (V - outs)
            if len(inter) == 1:
                u = inter.pop()
                flow[u] = v
                layer[u] = k
                outs.add(u)
                if u not in set(inputs):
                    past[u] = len(set(graph.neighbors(u)) & (V - outs))
                    if past[u] == 1:
                        nxt.add(u)
                for w in set(graph.neighbors(u)):
                    if past[w] > 0:
                        past[w] -= 1
                        if past[w] == 1:
                            nxt.add(w)
        if nxt:
            aux(nxt, k + 1)

    aux(C, 1)
    if len(flow) != len(nodes) - len(outputs):
        return None, None
    return flow, layer


def measurement_order(graph, inputs, layer):
    nodes = list(graph.nodes())
    mat = np.array([[1 if layer[i] > layer[j] else 0 for j in nodes] for i in nodes])
    sums = mat.sum(axis=1)
    groups = {}
    for i, s in enumerate(sums):
        groups.setdefault(s, []).append(nodes[i])
    order = [v for s in sorted(groups, reverse=True) for v in groups[s]]
    for i in list(inputs)[::-1]:
        order.remove(i)
        order.insert(0, i)
    return order


def trainable_nodes(graph, outputs):
    return [v for v in graph.nodes() if v not in outputs]


def grid_edges(n, m):
    edges = [(i * m + j, i * m + j + 1) for i in range(n) for j in range(m - 1)]
    edges += [(i * m + j, (i + 1) * m + j) for i in range(n - 1) for j in range(m)]
    return edges


def grid_order_closed_form(n, m):
    return [i * m + j for j in range(m) for i in range(n)]
