"""Unshielded-collider (v-structure) orientation on the host.

Same visiting order as gpucsl/pc/edge_orientation/v_structures.py:8-35, since
the order decides conflicting triples: ordered edges (v1, v2) of the skeleton
in column-major order (key v2 * p + v1), then the neighbours v3 of v2 in
ascending order.  v1 -> v2 <- v3 is oriented when v1 and v3 are not adjacent
and v2 is in neither sepset(v1, v3) nor sepset(v3, v1).
"""
import networkx as nx
import numpy as np


def orient_v_structure(dag: nx.DiGraph, separation_sets: np.ndarray, skeleton: nx.Graph = None) -> None:
    if skeleton is None:
        skeleton = dag.to_undirected()
    p = len(skeleton.nodes)

    def separates(v, a, b):
        return v in separation_sets[a][b]

    ordered = sorted(skeleton.to_directed().edges, key=lambda e: e[1] * p + e[0])
    for v1, v2 in ordered:
        for v3 in sorted(skeleton.neighbors(v2)):
            if v3 == v1 or skeleton.has_edge(v1, v3):
                continue
            if separates(v2, v1, v3) or separates(v2, v3, v1):
                continue
            dag.add_edges_from([(v1, v2), (v3, v2)])
            dag.remove_edges_from([(v2, v1), (v2, v3)])
