"""CPDAG from skeleton and separation sets (gpucsl/pc/edge_orientation/edge_orientation.py:25-36).

Graphs whose nodes are the integers 0..p-1 (every skeleton the PC classes
produce) are oriented by the library's host routine ``gpucsl_orient_edges``
on dense adjacency bytes; other node labels go through the networkx
restatement in ``v_structures`` / ``orientation_rules``.  Both follow the
reference's visiting order, so they give the same graph.
"""
import ctypes as ct

import networkx as nx
import numpy as np

from gpucsl_amd.pc.edge_orientation.orientation_rules import apply_rules
from gpucsl_amd.pc.edge_orientation.v_structures import orient_v_structure
from gpucsl_amd.pc.helpers import timed


def _dense_labels(g) -> bool:
    n = g.number_of_nodes()
    return all(isinstance(v, (int, np.integer)) and 0 <= v < n for v in g.nodes)


def orient_dense(skeleton: np.ndarray, separation_sets: np.ndarray) -> np.ndarray:
    """Dense form: skeleton (p, p) 0/1 symmetric, separation_sets (p, p, L)
    int32 -1 padded -> dag (p, p) uint8 with dag[i, j] = arc i -> j."""
    from gpucsl_amd import _native as nat

    skel = np.ascontiguousarray(np.asarray(skeleton) != 0, dtype=np.uint8)
    p = skel.shape[0]
    sep = np.ascontiguousarray(np.asarray(separation_sets, dtype=np.int32).reshape(p, p, -1))
    L = sep.shape[2]
    dag = np.empty((p, p), np.uint8)
    passes = ct.c_int(0)
    rc = nat.load().gpucsl_orient_edges(skel.ctypes.data, p, sep.ctypes.data if L else None, L, dag.ctypes.data,
                                        ct.byref(passes))
    if rc:
        raise ValueError("orient_dense: bad arguments")
    return dag


def orient_arcs(skeleton: np.ndarray, separation_sets: np.ndarray):
    """orient_dense, and the CPDAG's arcs as an (m, 2) int32 array in
    row-major order (the order of np.nonzero(dag)), listed by the library
    from the skeleton's neighbour lists instead of a scan of the p x p bytes."""
    from gpucsl_amd import _native as nat

    skel = np.ascontiguousarray(np.asarray(skeleton) != 0).view(np.uint8)  # bool bytes are 0/1
    p = skel.shape[0]
    sep = np.ascontiguousarray(np.asarray(separation_sets, dtype=np.int32).reshape(p, p, -1))
    L = sep.shape[2]
    dag = np.empty((p, p), np.uint8)
    cap = int(np.count_nonzero(skel))  # every CPDAG arc is a skeleton arc
    arcs = np.empty((max(cap, 1), 2), np.int32)
    m = ct.c_int64(0)
    passes = ct.c_int(0)
    rc = nat.load().gpucsl_orient_edges_arcs(skel.ctypes.data, p, sep.ctypes.data if L else None, L,
                                             dag.ctypes.data, arcs.ctypes.data, cap, ct.byref(m), ct.byref(passes))
    if rc:
        raise ValueError(f"orient_arcs: bad arguments (code {rc})")
    return dag, arcs[: m.value]


def _digraph_of_arcs(p: int, arcs: np.ndarray) -> nx.DiGraph:
    out = nx.DiGraph()
    out.add_nodes_from(range(p))
    out.add_edges_from(arcs.tolist())
    return out


@timed
def orient_skeleton(skeleton: np.ndarray, separation_sets: np.ndarray) -> nx.DiGraph:
    """orient_edges for a skeleton given as its (p, p) adjacency array (what
    the skeleton routines return): the same CPDAG without building a networkx
    graph of the skeleton first."""
    _, arcs = orient_arcs(skeleton, separation_sets)
    return _digraph_of_arcs(np.asarray(skeleton).shape[0], arcs)


@timed
def orient_edges(skeleton: nx.Graph, separation_sets: np.ndarray) -> nx.DiGraph:
    """skeleton: undirected graph (or a DiGraph with both arcs per edge);
    separation_sets: (p, p, max_level) int32, -1 padded."""
    if _dense_labels(skeleton):
        p = skeleton.number_of_nodes()
        adj = nx.to_numpy_array(skeleton, nodelist=range(p), weight=None, dtype=np.uint8)
        adj = np.maximum(adj, adj.T)
        return _digraph_of_arcs(p, orient_arcs(adj, separation_sets)[1])
    dag = skeleton.to_directed()
    orient_v_structure(dag, separation_sets, skeleton)
    apply_rules(dag)
    return dag
