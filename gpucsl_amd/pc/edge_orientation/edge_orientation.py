"""CPDAG from skeleton and separation sets (gpucsl/pc/edge_orientation/edge_orientation.py:25-36)."""
import networkx as nx
import numpy as np

from gpucsl_amd.pc.edge_orientation.orientation_rules import apply_rules
from gpucsl_amd.pc.edge_orientation.v_structures import orient_v_structure
from gpucsl_amd.pc.helpers import timed


@timed
def orient_edges(skeleton: nx.Graph, separation_sets: np.ndarray) -> nx.DiGraph:
    """skeleton: undirected graph (or a DiGraph with both arcs per edge);
    separation_sets: (p, p, max_level) int32, -1 padded."""
    dag = skeleton.to_directed()
    orient_v_structure(dag, separation_sets, skeleton)
    apply_rules(dag)
    return dag
