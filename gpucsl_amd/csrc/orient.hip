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
// No host memory is allocated by the caller's buffers; the snapshot is a
// std::vector owned here.
#include <cstdint>
#include <cstring>
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
        for (int j = 0; j < p; ++j)
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

// one pass of R1, R2, R3; returns whether any rule fired
bool apply_rules_once(Graph &dag, std::vector<uint8_t> &snap, const Nbrs &nbr) {
    const int p = dag.p;
    const size_t pp = (size_t)p * p;
    bool changed = false;
    Graph g{p, snap.data()};
    auto non_adjacent = [&](int x, int y) { return !g.has(x, y) && !g.has(y, x); };

    // R1: v1 -> v2 - v3, v1 and v3 non-adjacent  =>  v2 -> v3
    std::memcpy(snap.data(), dag.a, pp);
    for (int v2 = 0; v2 < p; ++v2)
        for (int v1 : nbr[v2]) {
            if (!g.has(v1, v2) || g.has(v2, v1)) continue;
            for (int v3 : nbr[v2]) {
                if (v3 == v1 || !g.has(v2, v3)) continue;
                if (g.has(v3, v2) && non_adjacent(v1, v3) && dag.has(v2, v3) && dag.has(v3, v2)) {
                    dag.set(v3, v2, 0);
                    changed = true;
                }
            }
        }

    // R2: v1 -> v3 -> v2 and v1 - v2  =>  v1 -> v2
    std::memcpy(snap.data(), dag.a, pp);
    for (int v2 = 0; v2 < p; ++v2)
        for (int v1 : nbr[v2]) {
            if (!g.has(v1, v2) || !g.has(v2, v1)) continue;
            for (int v3 : nbr[v1]) {
                if (!g.has(v1, v3) || !g.has(v3, v2)) continue;
                if (!g.has(v3, v1) && !g.has(v2, v3)) {
                    dag.set(v1, v2, 1);
                    dag.set(v2, v1, 0);
                    changed = true;
                }
            }
        }

    // R3: v1 - v3 -> v2, v1 - v4 -> v2, v3 and v4 non-adjacent, v1 - v2  =>  v1 -> v2
    std::memcpy(snap.data(), dag.a, pp);
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
                        dag.set(v1, v2, 1);
                        dag.set(v2, v1, 0);
                        changed = true;
                    }
        }
    return changed;
}

}  // namespace

extern "C" int gpucsl_orient_edges(const uint8_t *skeleton, int p, const int32_t *sepsets, int max_level,
                                   uint8_t *dag, int *passes) {
    if (!skeleton || !dag || p < 0 || max_level < 0 || (max_level > 0 && !sepsets)) return 1;
    const size_t pp = (size_t)p * p;
    std::memcpy(dag, skeleton, pp);
    Graph g{p, dag};
    const Nbrs nbr = skeleton_neighbours(skeleton, p);
    orient_v_structures(skeleton, p, sepsets, max_level, nbr, g);
    std::vector<uint8_t> snap(pp);
    int n = 0;
    for (;;) {
        ++n;
        if (!apply_rules_once(g, snap, nbr)) break;
    }
    if (passes) *passes = n;
    return 0;
}
