"""Resource graph container (`mentpy/mbqc/states/graphstate.py:14-51`)."""
from __future__ import annotations

import networkx as nx

__all__ = ["GraphState"]


class GraphState(nx.Graph):
    """A graph state: an undirected graph whose vertices are qubits in |+>
    and whose edges are CZ gates."""

    def __repr__(self):
        return f"GraphState with {self.number_of_nodes()} nodes and {self.number_of_edges()} edges."

    def __len__(self):
        return self.number_of_nodes()

    def __eq__(self, other):
        return nx.is_isomorphic(self, other)

    __hash__ = nx.Graph.__hash__

    def index_mapping(self):
        return {v: i for i, v in enumerate(self.nodes())}
