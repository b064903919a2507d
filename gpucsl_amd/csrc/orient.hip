// Edge orientation (CPDAG) on the host: v-structures, then Meek's rules R1-R3
// until nothing changes.  Replaces gpucsl/pc/edge_orientation/
// (v_structures.py:8-35, orientation_rules.py:7-89, edge_orientation.py:25-36),
// which walks networkx graphs; this walks dense p x p adjacency bytes and is
// meant for the graph sizes the GPU skeleton produces (thousands of nodes).
//
// The visiting order is the reference's, because orientations that conflict
// are resolved by it:
//   - v-structures: directed skeleton edges (v1, v2) in column-major order
//     (v2 major, v1 minor), third nodes v3 ascending over the neighbours of v2;
//     each unshielded collider v1 -> v2 <- v3 with v2 in neither sepset(v1, v3)
//     nor sepset(v3, v1) sets both arcs into v2, overriding earlier choices;
//   - every rule pass reads a snapshot of the graph taken before that rule
//     (R1, then R2 on the graph after R1, then R3) and writes the live graph;
//     R1 changes an edge only if it is still undirected in the live graph;
//     R2 and R3 count as a change whenever their condition holds, as the
//     reference's loop does.
// The neighbour lists and the rules' write logs are owned here.
#include <cstdint>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "../../include/gpucsl_hip.h"

namespace {

struct Graph {
    int p;
    uint8_t *a;  // a[i * p + j] != 0: arc i -> j
    bool has(int i, int j) const { return a[(size_t)i * p + j] != 0; }
    void set(int i, int j, uint8_t v) { a[(size_t)i * p + j] = v; }
};

bool in_sepset(const int32_t *sep, int p, int L, int v, int a, int b) {
    const int32_t *s = sep + ((size_t)a * p + b) * L;
    for (int k = 0; k < L; ++k)
        if (s[k] == v) return true;
    return false;
}

// Ascending skeleton neighbours of every node.  Orienting only removes one
// arc of an edge, so every arc the rules look at is a skeleton edge: walking
// neighbour lists visits the same (v1, v2, v3) triples, in the same order, as
// walking every node index and skipping non-arcs.
using Nbrs = std::vector<std::vector<int>>;

Nbrs skeleton_neighbours(const uint8_t *skel, int p) {
    Nbrs nbr(p);
    for (int i = 0; i < p; ++i) {
        const uint8_t *row = skel + (size_t)i * p;
        int j = 0;
        // eight bytes at a time: a sparse row is mostly zero words
        for (; j + 8 <= p; j += 8) {
            uint64_t w;
            std::memcpy(&w, row + j, 8);
            if (!w) continue;
            for (int k = 0; k < 8; ++k)
                if (row[j + k]) nbr[i].push_back(j + k);
        }
        for (; j < p; ++j)
            if (row[j]) nbr[i].push_back(j);
    }
    return nbr;
}

void orient_v_structures(const uint8_t *skel, int p, const int32_t *sep, int L, const Nbrs &nbr, Graph &dag) {
    for (int v2 = 0; v2 < p; ++v2)
        for (int v1 : nbr[v2]) {  // skel is symmetric: v1 -> v2 is a skeleton arc
            for (int v3 : nbr[v2]) {
                if (v3 == v1 || skel[(size_t)v1 * p + v3]) continue;
                if (L > 0 && (in_sepset(sep, p, L, v2, v1, v3) || in_sepset(sep, p, L, v2, v3, v1))) continue;
                dag.set(v1, v2, 1);
                dag.set(v3, v2, 1);
                dag.set(v2, v1, 0);
                dag.set(v2, v3, 0);
            }
        }
}

// one pass of R1, R2, R3; returns whether any rule fired.  Each rule reads
// the graph as it was before the rule and writes the live graph; here the
// rule's writes are logged and applied in order once the rule has visited
// every triple, so the graph itself is the pre-rule snapshot while the rule
// runs (no copy of the p x p bytes).  R2 and R3 read the snapshot only.  R1
// also asks whether an edge is still undirected in the live graph; R1 only
// removes arcs, so that is the snapshot minus the arcs R1 removed so far.
bool apply_rules_once(Graph &dag, std::vector<size_t> &log, std::unordered_set<size_t> &removed, const Nbrs &nbr) {
    const int p = dag.p;
    const size_t P = (size_t)p;
    bool changed = false;
    const Graph &g = dag;
    auto non_adjacent = [&](int x, int y) { return !g.has(x, y) && !g.has(y, x); };
    auto live_has = [&](int x, int y) { return g.has(x, y) && !removed.count((size_t)x * P + y); };

    // R1: v1 -> v2 - v3, v1 and v3 non-adjacent  =>  v2 -> v3
    removed.clear();
    for (int v2 = 0; v2 < p; ++v2)
        for (int v1 : nbr[v2]) {
            if (!g.has(v1, v2) || g.has(v2, v1)) continue;
            for (int v3 : nbr[v2]) {
                if (v3 == v1 || !g.has(v2, v3)) continue;
                if (g.has(v3, v2) && non_adjacent(v1, v3) && live_has(v2, v3) && live_has(v3, v2)) {
                    removed.insert((size_t)v3 * P + v2);
                    changed = true;
                }
            }
        }
    for (const size_t k : removed) dag.a[k] = 0;

    // R2: v1 -> v3 -> v2 and v1 - v2  =>  v1 -> v2
    // (log entries: arc index, the arc set; its reverse cleared)
    log.clear();
    for (int v2 = 0; v2 < p; ++v2)
        for (int v1 : nbr[v2]) {
            if (!g.has(v1, v2) || !g.has(v2, v1)) continue;
            for (int v3 : nbr[v1]) {
                if (!g.has(v1, v3) || !g.has(v3, v2)) continue;
                if (!g.has(v3, v1) && !g.has(v2, v3)) {
                    log.push_back((size_t)v1 * P + v2);
                    changed = true;
                }
            }
        }
    for (const size_t k : log) {
        dag.a[k] = 1;
        dag.a[(k % P) * P + k / P] = 0;
    }

    // R3: v1 - v3 -> v2, v1 - v4 -> v2, v3 and v4 non-adjacent, v1 - v2  =>  v1 -> v2
    log.clear();
    std::vector<int> C;
    for (int v2 = 0; v2 < p; ++v2)
        for (int v1 : nbr[v2]) {
            if (!g.has(v1, v2) || !g.has(v2, v1)) continue;
            C.clear();
            for (int v : nbr[v1])
                if (g.has(v1, v) && g.has(v, v1) && g.has(v, v2) && !g.has(v2, v)) C.push_back(v);
            for (size_t x = 0; x < C.size(); ++x)
                for (size_t y = x + 1; y < C.size(); ++y)
                    if (non_adjacent(C[x], C[y])) {
                        log.push_back((size_t)v1 * P + v2);
                        changed = true;
                    }
        }
    for (const size_t k : log) {
        dag.a[k] = 1;
        dag.a[(k % P) * P + k / P] = 0;
    }
    return changed;
}

int orient(const uint8_t *skeleton, int p, const int32_t *sepsets, int max_level, uint8_t *dag, Nbrs &nbr,
           int *passes) {
    if (!skeleton || !dag || p < 0 || max_level < 0 || (max_level > 0 && !sepsets)) return 1;
    const size_t pp = (size_t)p * p;
    std::memcpy(dag, skeleton, pp);
    Graph g{p, dag};
    nbr = skeleton_neighbours(skeleton, p);
    orient_v_structures(skeleton, p, sepsets, max_level, nbr, g);
    std::vector<size_t> log;
    std::unordered_set<size_t> removed;
    int n = 0;
    for (;;) {
        ++n;
        if (!apply_rules_once(g, log, removed, nbr)) break;
    }
    if (passes) *passes = n;
    return 0;
}

}  // namespace

extern "C" int gpucsl_orient_edges(const uint8_t *skeleton, int p, const int32_t *sepsets, int max_level,
                                   uint8_t *dag, int *passes) {
    Nbrs nbr;
    return orient(skeleton, p, sepsets, max_level, dag, nbr, passes);
}

extern "C" int gpucsl_orient_edges_arcs(const uint8_t *skeleton, int p, const int32_t *sepsets, int max_level,
                                        uint8_t *dag, int32_t *arcs, int64_t arcs_cap, int64_t *n_arcs,
                                        int *passes) {
    if (!n_arcs || arcs_cap < 0 || (arcs_cap > 0 && !arcs)) return 1;
    Nbrs nbr;
    const int rc = orient(skeleton, p, sepsets, max_level, dag, nbr, passes);
    if (rc) return rc;
    // every arc of the CPDAG is a skeleton arc: walk the neighbour lists
    int64_t m = 0;
    for (int i = 0; i < p; ++i)
        for (int j : nbr[i])
            if (dag[(size_t)i * p + j]) {
                if (m < arcs_cap) {
                    arcs[2 * m] = i;
                    arcs[2 * m + 1] = j;
                }
                ++m;
            }
    *n_arcs = m;
    return m > arcs_cap ? 2 : 0;
}
