"""Meek's rules R1-R3 applied to a fixpoint on the host.

Same semantics and visiting order as
gpucsl/pc/edge_orientation/orientation_rules.py:7-89: each rule scans a copy of
the graph taken before the rule (edges in column-major order, candidate third
vertices ascending) and writes its orientations into the live graph.  An
undirected edge is a pair of opposite arcs.
"""
import itertools

import networkx as nx


def _column_major(p):
    return lambda e: e[1] * p + e[0]


def _adjacent(g, a, b):
    return g.has_edge(a, b) or g.has_edge(b, a)


def _orient(dag, a, b):
    """a - b becomes a -> b."""
    dag.add_edge(a, b)
    dag.remove_edges_from([(b, a)])


def _rule1(dag, key):
    # a -> b - c with a, c non-adjacent: b -> c
    snap = dag.copy()
    changed = False
    for a, b in sorted(snap.edges, key=key):
        if snap.has_edge(b, a):
            continue
        for c in sorted(snap.successors(b)):
            if c == a or not snap.has_edge(c, b) or _adjacent(snap, a, c):
                continue
            if dag.has_edge(b, c) and dag.has_edge(c, b):
                _orient(dag, b, c)
                changed = True
    return changed


def _rule2(dag, key):
    # a -> c -> b with a - b: a -> b
    snap = dag.copy()
    changed = False
    for a, b in sorted(snap.edges, key=key):
        if not snap.has_edge(b, a):
            continue
        for c in sorted(set(snap.successors(a)) & set(snap.predecessors(b))):
            if not snap.has_edge(c, a) and not snap.has_edge(b, c):
                _orient(dag, a, b)
                changed = True
    return changed


def _rule3(dag, key):
    # a - c -> b, a - d -> b, a - b, c and d non-adjacent: a -> b
    snap = dag.copy()
    changed = False
    for a, b in sorted(snap.edges, key=key):
        if not snap.has_edge(b, a):
            continue
        undirected_a = set(snap.successors(a)) & set(snap.predecessors(a))
        into_b = set(snap.predecessors(b)) - set(snap.successors(b))
        for c, d in itertools.combinations(sorted(undirected_a & into_b), 2):
            if not _adjacent(snap, c, d):
                _orient(dag, a, b)
                changed = True
    return changed


def apply_rules(dag: nx.DiGraph):
    key = _column_major(len(dag.nodes))
    while True:
        changed = _rule1(dag, key)
        changed = _rule2(dag, key) or changed
        changed = _rule3(dag, key) or changed
        if not changed:
            return
