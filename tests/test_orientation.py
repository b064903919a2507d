"""Edge orientation (CPU): Meek rules R1-R3 and v-structures on the hand-built
graphs of the reference's tests (tests/edge_orientation/test_orientation_rules.py:16-72,
tests/edge_orientation/test_v_structures.py:37-116), plus CPDAG invariants."""
import networkx as nx
import numpy as np

from gpucsl_amd.pc.edge_orientation.edge_orientation import orient_edges
from gpucsl_amd.pc.edge_orientation.orientation_rules import apply_rules
from gpucsl_amd.pc.edge_orientation.v_structures import orient_v_structure


def sepsets(p, L, entries):
    s = np.full((p, p, L), -1, np.int32)
    for (a, b), vals in entries.items():
        s[a, b, : len(vals)] = vals
    return s


def digraph(nodes, arcs):
    g = nx.DiGraph()
    g.add_nodes_from(nodes)
    g.add_edges_from(arcs)
    return g


def test_rule1_chain_continues_direction():
    # 1 -> 2 - 3 with 1, 3 non-adjacent: 2 -> 3
    g = digraph([1, 2, 3], [(1, 2), (2, 3), (3, 2)])
    apply_rules(g)
    assert g.has_edge(1, 2) and g.has_edge(2, 3) and not g.has_edge(3, 2)


def test_rule2_avoids_cycle():
    # 1 - 3 undirected and the chain 1 -> 2 -> 3: R2 orients 1 -> 3, since
    # 3 -> 1 would close the cycle 1 -> 2 -> 3 -> 1
    g = digraph([1, 2, 3], [(1, 3), (3, 1), (1, 2), (2, 3)])
    apply_rules(g)
    assert not g.has_edge(3, 1)
    assert g.has_edge(1, 3) and g.has_edge(1, 2) and g.has_edge(2, 3)


def test_rule3_two_collider_parents():
    g = digraph([1, 2, 3, 4], [(1, 2), (2, 1), (1, 3), (3, 1), (1, 4), (4, 1), (3, 2), (4, 2)])
    apply_rules(g)
    assert not g.has_edge(2, 1)
    for e in [(1, 2), (1, 3), (3, 1), (1, 4), (4, 1), (3, 2), (4, 2)]:
        assert g.has_edge(*e)


def test_v_structure_respects_sepset():
    edges = [(1, 2), (2, 1), (2, 3), (3, 2)]
    g = digraph([1, 2, 3], edges)
    orient_v_structure(g, sepsets(4, 2, {(1, 3): [2], (3, 1): [2]}))
    assert all(g.has_edge(*e) for e in edges)             # 2 separates 1 and 3: no collider
    orient_v_structure(g, sepsets(4, 2, {}))
    assert g.has_edge(1, 2) and g.has_edge(3, 2)
    assert not g.has_edge(2, 1) and not g.has_edge(2, 3)


def test_v_structure_conflicting_triples_follow_visit_order():
    # path 1 - 2 - 3 - 4 with empty sepsets: colliders at 2 and 3 conflict; the
    # column-major visit order resolves it as 1 -> 2 -> 3 <- 4
    g = digraph([1, 2, 3, 4], [(1, 2), (2, 1), (2, 3), (3, 2), (3, 4), (4, 3)])
    orient_v_structure(g, sepsets(5, 2, {}))
    assert g.has_edge(1, 2) and g.has_edge(2, 3) and g.has_edge(4, 3)
    assert not g.has_edge(2, 1) and not g.has_edge(3, 2) and not g.has_edge(3, 4)


def test_v_structure_partially_directed_input():
    g = digraph([1, 2, 3], [(1, 2), (2, 1), (3, 2)])
    orient_v_structure(g, sepsets(4, 2, {(1, 3): [2], (3, 1): [2]}))
    assert g.has_edge(1, 2) and g.has_edge(2, 1) and g.has_edge(3, 2)
    orient_v_structure(g, sepsets(4, 2, {}))
    assert not g.has_edge(2, 1) and g.has_edge(1, 2)


def test_orient_edges_collider_and_chain_cpdag():
    # skeleton 0 - 2 - 1 (collider at 2, 0 and 1 separated by the empty set) and 2 - 3
    skel = np.zeros((4, 4), np.uint16)
    for a, b in [(0, 2), (1, 2), (2, 3)]:
        skel[a, b] = skel[b, a] = 1
    sep = sepsets(4, 2, {})
    sep[0, 3, 0] = sep[3, 0, 0] = 2
    sep[1, 3, 0] = sep[3, 1, 0] = 2
    dag, runtime = orient_edges(nx.DiGraph(skel), sep)
    assert runtime >= 0
    assert set(dag.edges) == {(0, 2), (1, 2), (2, 3)}     # R1 orients 2 -> 3


def test_orient_edges_is_a_pdag_of_the_skeleton():
    rng = np.random.default_rng(0)
    p = 12
    skel = np.triu((rng.random((p, p)) < 0.3).astype(np.uint16), 1)
    skel = skel + skel.T
    sep = np.full((p, p, 2), -1, np.int32)
    dag, _ = orient_edges(nx.DiGraph(skel), sep)
    und = {frozenset(e) for e in dag.edges}
    want = {frozenset((a, b)) for a in range(p) for b in range(p) if skel[a, b]}
    assert und == want                                    # no edge added or lost


def _python_orientation(skel, sep):
    g = nx.DiGraph(skel)
    dag = g.to_directed()
    orient_v_structure(dag, sep, g)
    apply_rules(dag)
    return set(dag.edges)


def test_native_orientation_equals_the_reference_order_restatement():
    """The library's dense routine against the networkx restatement of the
    reference's visiting order, on random skeletons with random sepsets."""
    from gpucsl_amd.pc.edge_orientation.edge_orientation import orient_dense

    rng = np.random.default_rng(7)
    for trial in range(60):
        p = int(rng.integers(3, 40))
        dens = rng.uniform(0.05, 0.5)
        skel = np.triu((rng.random((p, p)) < dens).astype(np.uint16), 1)
        skel = skel + skel.T
        L = int(rng.integers(0, 4))
        sep = np.full((p, p, L), -1, np.int32)
        for a in range(p):
            for b in range(a + 1, p):
                if not skel[a, b] and L and rng.random() < 0.7:
                    k = int(rng.integers(1, L + 1))
                    vals = np.sort(rng.choice(p, size=k, replace=False))
                    sep[a, b, :k] = vals
                    sep[b, a, :k] = vals
        want = _python_orientation(skel, sep)
        dag = orient_dense(skel, sep)
        got = {(int(i), int(j)) for i, j in zip(*np.nonzero(dag))}
        assert got == want, f"trial {trial} p={p}"


def test_arc_list_is_the_dense_result_in_row_major_order():
    """orient_arcs lists exactly the nonzeros of its dense CPDAG, in the order
    of np.nonzero, and equals the networkx restatement; also for a skeleton
    with no edges and for a CPDAG with every skeleton edge oriented."""
    from gpucsl_amd.pc.edge_orientation.edge_orientation import orient_arcs

    rng = np.random.default_rng(11)
    for trial in range(30):
        p = int(rng.integers(1, 60))
        skel = np.triu((rng.random((p, p)) < rng.uniform(0.0, 0.4)).astype(np.uint16), 1)
        skel = skel + skel.T
        sep = np.full((p, p, 2), -1, np.int32)
        dag, arcs = orient_arcs(skel, sep)
        ii, jj = np.nonzero(dag)
        assert arcs.dtype == np.int32 and arcs.shape == (len(ii), 2), f"trial {trial}"
        assert np.array_equal(arcs[:, 0], ii) and np.array_equal(arcs[:, 1], jj), f"trial {trial}"
        assert {(int(a), int(b)) for a, b in arcs} == _python_orientation(skel, sep), f"trial {trial}"
    # a collider 0 -> 2 <- 1 (0 and 1 separated by the empty set): both arcs oriented
    skel = np.zeros((3, 3), np.uint16)
    skel[0, 2] = skel[2, 0] = skel[1, 2] = skel[2, 1] = 1
    dag, arcs = orient_arcs(skel, np.full((3, 3, 1), -1, np.int32))
    assert arcs.tolist() == [[0, 2], [1, 2]]


def test_arc_list_capacity_contract():
    """gpucsl_orient_edges_arcs: a buffer too small for the arcs returns 2
    with the full count and a complete dense result; a null count pointer or
    a negative capacity is a bad argument (1)."""
    import ctypes as ct

    from gpucsl_amd import _native as nat
    from gpucsl_amd.pc.edge_orientation.edge_orientation import orient_arcs

    rng = np.random.default_rng(5)
    p = 25
    skel = np.triu((rng.random((p, p)) < 0.3).astype(np.uint8), 1)
    skel = np.ascontiguousarray(skel + skel.T)
    sep = np.full((p, p, 1), -1, np.int32)
    want_dag, want_arcs = orient_arcs(skel, sep)
    assert len(want_arcs) > 2
    lib = nat.load()
    dag = np.zeros((p, p), np.uint8)
    arcs = np.full((2, 2), -7, np.int32)
    m = ct.c_int64(-1)
    rc = lib.gpucsl_orient_edges_arcs(skel.ctypes.data, p, sep.ctypes.data, 1, dag.ctypes.data, arcs.ctypes.data, 2,
                                      ct.byref(m), None)
    assert rc == 2 and m.value == len(want_arcs)
    assert np.array_equal(dag, want_dag) and np.array_equal(arcs, want_arcs[:2])
    assert lib.gpucsl_orient_edges_arcs(skel.ctypes.data, p, sep.ctypes.data, 1, dag.ctypes.data,
                                        arcs.ctypes.data, 2, None, None) == 1
    assert lib.gpucsl_orient_edges_arcs(skel.ctypes.data, p, sep.ctypes.data, 1, dag.ctypes.data,
                                        arcs.ctypes.data, -1, ct.byref(m), None) == 1


def test_orient_edges_uses_native_and_matches_on_pc_output():
    rng = np.random.default_rng(3)
    p = 30
    skel = np.triu((rng.random((p, p)) < 0.15).astype(np.uint16), 1)
    skel = skel + skel.T
    sep = np.full((p, p, 2), -1, np.int32)
    dag, _ = orient_edges(nx.DiGraph(skel), sep)
    assert set(dag.edges) == _python_orientation(skel, sep)
    assert sorted(dag.nodes) == list(range(p))


def _dsep_sepsets(dag_arcs, p, L):
    """Skeleton and separation sets of a DAG by d-separation: i and j are
    adjacent iff no subset of the other nodes separates them; the recorded set
    is the first separating set in (size, lexicographic) order."""
    import itertools

    g = nx.DiGraph()
    g.add_nodes_from(range(p))
    g.add_edges_from(dag_arcs)
    skel = np.zeros((p, p), np.uint16)
    sep = np.full((p, p, L), -1, np.int32)
    for a in range(p):
        for b in range(a + 1, p):
            rest = [v for v in range(p) if v not in (a, b)]
            found = None
            for k in range(0, len(rest) + 1):
                for S in itertools.combinations(rest, k):
                    if nx.is_d_separator(g, {a}, {b}, set(S)):
                        found = S
                        break
                if found is not None:
                    break
            if found is None:
                skel[a, b] = skel[b, a] = 1
            else:
                assert len(found) <= L
                sep[a, b, : len(found)] = found
                sep[b, a, : len(found)] = found
    return skel, sep


def _cpdag_by_enumeration(skel, dag_arcs):
    """The CPDAG by definition: every acyclic orientation of the skeleton with
    the same v-structures is Markov equivalent (Verma and Pearl); an arc is
    directed in the CPDAG iff every one of them has it."""
    import itertools

    p = skel.shape[0]
    edges = [(a, b) for a in range(p) for b in range(a + 1, p) if skel[a, b]]

    def vstructs(arcs):
        par = {v: {u for (u, w) in arcs if w == v} for v in range(p)}
        return {(min(u, w), v, max(u, w)) for v in range(p) for u in par[v] for w in par[v]
                if u < w and not skel[u, w]}

    want = vstructs(set(dag_arcs))
    members = []
    for bits in itertools.product((0, 1), repeat=len(edges)):
        arcs = {(a, b) if t == 0 else (b, a) for (a, b), t in zip(edges, bits)}
        g = nx.DiGraph(list(arcs))
        g.add_nodes_from(range(p))
        if nx.is_directed_acyclic_graph(g) and vstructs(arcs) == want:
            members.append(arcs)
    assert members
    out = set()
    for a, b in edges:
        fwd = all((a, b) in m for m in members)
        bwd = all((b, a) in m for m in members)
        if fwd:
            out.add((a, b))
        elif bwd:
            out.add((b, a))
        else:
            out |= {(a, b), (b, a)}
    return out


def test_orientation_recovers_the_cpdag_of_random_dags():
    """Given a DAG's skeleton and separating sets (d-separation), v-structures
    followed by Meek's R1-R3 to a fixpoint give the DAG's CPDAG (Meek 1995).
    The CPDAG is computed independently, by enumerating the Markov
    equivalence class.  Both the library's dense routine and the networkx
    restatement are checked."""
    rng = np.random.default_rng(11)
    checked_directed = 0
    for trial in range(40):
        p = int(rng.integers(4, 8))
        order = rng.permutation(p)
        arcs = [(int(order[a]), int(order[b])) for a in range(p) for b in range(a + 1, p)
                if rng.random() < 0.4]
        skel, sep = _dsep_sepsets(arcs, p, p - 2)
        want = _cpdag_by_enumeration(skel, arcs)
        dag, _ = orient_edges(nx.DiGraph(skel), sep)
        assert set(dag.edges) == want, f"trial {trial}: {arcs}"
        assert _python_orientation(skel, sep) == want, f"trial {trial}"
        checked_directed += sum(1 for (a, b) in want if (b, a) not in want)
    assert checked_directed > 20


def test_orient_skeleton_array_equals_orient_edges():
    """The array form PC.execute uses gives the graph orient_edges gives for
    the networkx skeleton."""
    from gpucsl_amd.pc.edge_orientation.edge_orientation import orient_skeleton

    rng = np.random.default_rng(5)
    for trial in range(20):
        p = int(rng.integers(3, 30))
        skel = np.triu((rng.random((p, p)) < 0.3).astype(np.uint16), 1)
        skel = skel + skel.T
        sep = np.full((p, p, 2), -1, np.int32)
        for a in range(p):
            for b in range(a + 1, p):
                if not skel[a, b] and rng.random() < 0.5:
                    sep[a, b, 0] = sep[b, a, 0] = int(rng.integers(0, p))
        a, _ = orient_skeleton(skel, sep)
        b, _ = orient_edges(nx.DiGraph(skel), sep)
        assert set(a.edges) == set(b.edges) and sorted(a.nodes) == sorted(b.nodes)
