"""Measurement-pattern description: resource graph, I/O nodes, measurements, order.

``MBQCircuit`` keeps the constructor and attribute semantics of mentpy's class
(`mentpy/mbqc/mbqcircuit.py:56-157, 195-373, 390-422`): node relabelling to
``0..N-1``, the default trainable ``Ment("XY")`` on every non-output node,
``trainable_nodes`` in measurement-dict order, the causal flow and the
measurement order derived from it (inputs first).  The HIP backends read only
the attributes listed in ``mentpy_amd.simulators.hip_simulator._pattern_arrays``.
"""
from __future__ import annotations

import copy
from functools import cached_property
from typing import Callable, Dict, List, Optional

import networkx as nx

from ..operators.ment import Ment
from .flow import Flow, check_if_flow, find_cflow
from .graphstate import GraphState

__all__ = ["MBQCircuit", "hstack", "vstack"]


class MBQCircuit:
    """A measurement pattern on a graph state.

    Parameters mirror mentpy: ``graph``, ``input_nodes``, ``output_nodes``,
    ``measurements`` (node -> ``Ment`` or ``None``), ``default_measurement``,
    ``flow``/``partial_order`` (both or neither), ``measurement_order``,
    ``relabel_indices``.
    """

    def __init__(self, graph: GraphState, input_nodes: List[int] = (),
                 output_nodes: List[int] = (), measurements: Optional[Dict] = None,
                 default_measurement: Optional[Ment] = None,
                 flow: Optional[Callable] = None, partial_order: Optional[Callable] = None,
                 measurement_order: Optional[List[int]] = None,
                 relabel_indices: bool = True) -> None:
        input_nodes = list(input_nodes)
        output_nodes = list(output_nodes)
        if default_measurement is None:
            default_measurement = Ment("XY")
        if relabel_indices:
            mapping = dict(zip(sorted(graph.nodes), range(graph.number_of_nodes())))
            inverse = {v: k for k, v in mapping.items()}
            graph = nx.relabel_nodes(graph, mapping)
            input_nodes = [mapping[i] for i in input_nodes]
            output_nodes = [mapping[i] for i in output_nodes]
            if flow is not None:
                user_flow = flow
                flow = lambda x: mapping[user_flow(inverse[x])]
            if partial_order is not None:
                user_order = partial_order
                partial_order = lambda x, y: user_order(inverse[x], inverse[y])
            if measurement_order is not None:
                measurement_order = [mapping[i] for i in measurement_order]
            if measurements is not None:
                measurements = {mapping[k]: v for k, v in measurements.items()}
        if not isinstance(graph, GraphState):
            graph = GraphState(graph)
        self._graph = graph

        missing_in = [v for v in input_nodes if v not in graph.nodes]
        if missing_in:
            raise ValueError(f"Input nodes {input_nodes} are not in the graph. "
                             f"Graph nodes are {graph.nodes}")
        if any(v not in graph.nodes for v in output_nodes):
            raise ValueError(f"Output nodes {output_nodes} are not in the graph. "
                             f"Graph nodes are {graph.nodes}")
        self._input_nodes = input_nodes
        self._output_nodes = output_nodes
        if not isinstance(default_measurement, Ment):
            raise ValueError(f"Default measurement {default_measurement} is not an instance of Ment.")
        self._default_measurement = default_measurement

        if measurements is None:
            measurements = {node: default_measurement for node in self.outputc}
            for node in self.output_nodes:
                measurements[node] = None
        else:
            measurements = dict(measurements)
            bad = [v for v in measurements if v not in graph.nodes]
            if bad:
                raise ValueError(f"Nodes {bad} are not in the graph.")
            if not all(isinstance(v, Ment) or v is None for v in measurements.values()):
                raise ValueError(f"Values {measurements.values()} are not instances of Ment.")
            for node in graph.nodes:
                if node not in measurements:
                    measurements[node] = default_measurement if node in self.outputc else None
        self._measurements = measurements
        self._flow, self._partial_order = None, None
        self._update_attributes()

        if flow is None or partial_order is None:
            flow, partial_order, _, _ = find_cflow(graph, input_nodes, output_nodes)
            self.gflow = Flow(graph, input_nodes, output_nodes)
        else:
            check_if_flow(graph, input_nodes, output_nodes, flow, partial_order)
            self.gflow = None
        self._flow = flow
        self._partial_order = partial_order
        if measurement_order is None and flow is not None:
            measurement_order = self.calculate_order()
        self._quantum_output_nodes = [n for n, m in self._measurements.items() if m is None]
        self._measurement_order = measurement_order

    def __repr__(self) -> str:
        return f"MBQCircuit with {self.graph.number_of_nodes()} qubits."

    def __len__(self) -> int:
        return len(self.graph)

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        try:
            return getattr(self._graph, name)
        except AttributeError:
            pass
        gflow = self.__dict__.get("gflow")
        if gflow is not None and hasattr(gflow, name):
            return getattr(gflow, name)
        raise AttributeError(f"Attribute {name} not found in MBQCircuit.")

    def __setitem__(self, key, value):
        if key not in self.graph.nodes:
            raise ValueError(f"Node {key} is not in the graph.")
        if not isinstance(value, Ment):
            raise ValueError(f"Value {value} is not a Measurement object.")
        self._measurements[key] = value
        self._update_attributes()

    def __getitem__(self, key):
        try:
            return self._measurements[key]
        except KeyError:
            raise ValueError(f"Node {key} is not in the graph.")

    def __delitem__(self, key):
        if key not in self.graph.nodes:
            raise ValueError(f"Node {key} is not in the graph.")
        self._measurements[key] = None

    # ----- attributes -------------------------------------------------------
    @property
    def measurements(self) -> Dict[int, Ment]:
        return self._measurements

    @measurements.setter
    def measurements(self, measurements):
        if not all(v in self.graph.nodes for v in measurements.keys()):
            raise ValueError(f"Nodes {measurements.keys()} are not in the graph.")
        if not all(isinstance(v, Ment) for v in measurements.values()):
            raise ValueError(f"Values {measurements.values()} are not Measurement objects.")
        self._measurements = measurements
        self._update_attributes()

    @property
    def graph(self) -> GraphState:
        return self._graph

    @property
    def input_nodes(self) -> List[int]:
        return self._input_nodes

    @property
    def output_nodes(self) -> List[int]:
        return self._output_nodes

    @property
    def quantum_output_nodes(self) -> List[int]:
        return self._quantum_output_nodes

    @property
    def classical_output_nodes(self) -> List[int]:
        return self._classical_output_nodes

    @property
    def trainable_nodes(self) -> List[int]:
        return self._trainable_nodes

    @trainable_nodes.setter
    def trainable_nodes(self, nodes):
        if not all(v in self.graph.nodes for v in nodes):
            raise ValueError(f"Trainable nodes {nodes} are not in the graph. "
                             f"Graph nodes are {self.graph.nodes}")
        self._trainable_nodes = list(nodes)

    @property
    def planes(self) -> dict:
        return self._planes

    @property
    def flow(self) -> Callable:
        return self._flow

    @property
    def partial_order(self) -> Callable:
        return self._partial_order

    @property
    def depth(self):
        return self.gflow.depth if self.gflow is not None else None

    @property
    def measurement_order(self) -> List[int]:
        return self._measurement_order

    @measurement_order.setter
    def measurement_order(self, order):
        order = list(order)
        for i in range(len(order)):
            for j in range(i + 1, len(order)):
                if self.partial_order(order[i], order[j]):
                    raise ValueError(f"Invalid measurement order {order}.")
        self._measurement_order = order

    @cached_property
    def outputc(self) -> List:
        return [v for v in self.graph.nodes() if v not in self.output_nodes]

    @cached_property
    def inputc(self) -> List:
        return [v for v in self.graph.nodes() if v not in self.input_nodes]

    def ordered_layers(self, train_indices=False):
        if self.gflow is None or self.gflow.func is None:
            return None
        if train_indices:
            return [[self.trainable_nodes.index(n) for n in layer]
                    for layer in self.gflow.layers[:-1]]
        return self.gflow.layers

    def _update_attributes(self) -> None:
        trainable, planes, q_out, c_out = [], {}, [], []
        for node, ment in self._measurements.items():
            if ment is not None:
                if ment.is_trainable():
                    trainable.append(node)
                planes[node] = ment.plane
                self._measurements[node] = copy.deepcopy(ment)
                self._measurements[node].node_id = node
                if node in self._output_nodes:
                    c_out.append(node)
            else:
                planes[node] = ""
                if node in self._output_nodes:
                    q_out.append(node)
        self._trainable_nodes = trainable
        self._planes = planes
        self._quantum_output_nodes = q_out
        self._classical_output_nodes = c_out

    def calculate_order(self) -> List[int]:
        """Measurement order: nodes grouped by how many nodes they precede.

        Equivalent to `mentpy/mbqc/mbqcircuit.py:390-422`: for every node count
        the nodes it precedes in the partial order, group by that count in
        decreasing order (ties kept in graph order), then move the inputs to
        the front in ``input_nodes`` order.
        """
        nodes = list(self.graph.nodes())
        counts = [sum(1 for j in nodes if self.partial_order(i, j)) for i in nodes]
        groups: Dict[int, List[int]] = {}
        for node, c in zip(nodes, counts):
            groups.setdefault(c, []).append(node)
        order = [n for c in sorted(groups, reverse=True) for n in groups[c]]
        for i in self.input_nodes[::-1]:
            order.remove(i)
            order.insert(0, i)
        return order

    def add_edge(self, u, v):
        self.graph.add_edge(u, v)
        try:
            self.__init__(self.graph, self.input_nodes, self.output_nodes)
        except Exception as e:
            self.graph.remove_edge(u, v)
            raise ValueError(f"Cannot add edge between {u} and {v}.\n" + str(e))

    def add_edges_from(self, edges, **kwargs):
        new_graph = self.graph.copy()
        new_graph.add_edges_from(edges, **kwargs)
        self.__init__(new_graph, self.input_nodes, self.output_nodes)


def _hstack2(a: MBQCircuit, b: MBQCircuit) -> MBQCircuit:
    """Wire the outputs of ``a`` into the inputs of ``b`` by edge contraction
    (`mentpy/mbqc/mbqcircuit.py:561-610`)."""
    if len(a.output_nodes) != len(b.input_nodes):
        raise ValueError("The output of the first state must be the same size as the "
                         "input of the second state.")
    graph = nx.disjoint_union(a.graph, b.graph)
    nodes1, nodes2 = list(a.graph.nodes()), list(b.graph.nodes())
    n1 = len(nodes1)
    for i in range(len(a.output_nodes)):
        graph.add_edge(nodes1.index(a.output_nodes[i]), nodes2.index(b.input_nodes[i]) + n1)
    inputs = [nodes1.index(i) for i in a.input_nodes]
    outputs = [nodes2.index(j) + n1 for j in b.output_nodes]
    meas = {nodes1.index(i): m for i, m in a.measurements.items()}
    meas.update({nodes2.index(i) + n1: m for i, m in b.measurements.items()})
    for i, j in zip(a.output_nodes, b.input_nodes):
        graph.add_edge(i, j + len(a.graph))
        graph = nx.contracted_edge(graph, (j + len(a.graph), i), self_loops=False)
        del meas[i]
    return MBQCircuit(GraphState(graph), inputs, outputs, measurements=meas)


def _vstack2(a: MBQCircuit, b: MBQCircuit) -> MBQCircuit:
    """Disjoint union of two patterns (`mentpy/mbqc/mbqcircuit.py:527-558`)."""
    graph = nx.disjoint_union(a.graph, b.graph)
    n = len(a.graph)
    meas = dict(a.measurements)
    meas.update({i + n: m for i, m in b.measurements.items()})
    return MBQCircuit(GraphState(graph), a.input_nodes + [i + n for i in b.input_nodes],
                      a.output_nodes + [i + n for i in b.output_nodes], measurements=meas)


def merge(state1: MBQCircuit, state2: MBQCircuit, along=()) -> MBQCircuit:
    """Glue ``state2`` onto ``state1`` by identifying output ``i`` of the first
    with input ``j`` of the second for each ``(i, j)`` in ``along``
    (`mentpy/mbqc/mbqcircuit.py:459-497`): the merged node keeps the second
    pattern's label (shifted by ``len(state1.graph)``) and its measurement; the
    remaining inputs of the second pattern become inputs, and each merged
    output of the first is replaced by the matching output of the second."""
    along = list(along)
    for i, j in along:
        if i not in state1.output_nodes or j not in state2.input_nodes:
            raise ValueError(f"Cannot merge states at indices {i} and {j}")
    n1 = len(state1.graph)
    graph = nx.disjoint_union(state1.graph, state2.graph)
    along1, along2 = zip(*along) if along else ((), ())
    inputs = list(state1.input_nodes) + [i + n1 for i in state2.input_nodes if i not in along2]
    outputs, used = [], []
    for j in state1.output_nodes:
        if j in along1:
            k = list(state2.input_nodes).index(along2[along1.index(j)])
            used.append(k)
            outputs.append(state2.output_nodes[k] + n1)
        else:
            outputs.append(j)
    outputs += [i + n1 for k, i in enumerate(state2.output_nodes) if k not in used]
    meas = dict(state1.measurements)
    meas.update({i + n1: m for i, m in state2.measurements.items()})
    for i, j in along:
        graph.add_edge(i, j + n1)
        graph = nx.contracted_edge(graph, (j + n1, i), self_loops=False)
        del meas[i]
    return MBQCircuit(GraphState(graph), inputs, outputs, measurements=meas)


def hstack(states):
    from functools import reduce
    if len(states) == 0:
        raise ValueError("Cannot horizontally stack an empty list of states.")
    return states[0] if len(states) == 1 else reduce(_hstack2, states)


def vstack(states):
    from functools import reduce
    if len(states) == 0:
        raise ValueError("Cannot vertically stack an empty list of states.")
    return states[0] if len(states) == 1 else reduce(_vstack2, states)
