// Host-side preprocessing for the de Pina support-vector phase loop.
//
// Every step whose output order decides which cycle a phase selects is reproduced with the
// same standard-library containers and the same insertion/sort sequences as the reference,
// because libstdc++'s std::sort (unstable), std::priority_queue pop order among equal keys and
// std::unordered_map iteration order are all functions of exactly those sequences:
//   O1 CSR edge order          include/CsrGraph.h:113-174, include/CsrGraphMulti.h:96-168
//   O2 relabel order           include/CsrGraphMulti.h:229-300
//   O3 spanning-tree root      include/CsrGraphDegreeTwoRemoval.cpp:103-112
//   O4 coordinate order        include/CsrGraphMultiSpanningTree.cpp:11-135
//   O5 FVS reverse-delete      include/FVS.cpp:166-227
//   O6 Dijkstra pop order      include/Dijkstra.h:78-127
//   O7 SP non-tree edge order  include/Dijkstra.h:129-149
//   O8 candidate rank          mcb/mcb_cycle_deg_two_remove.cpp:244-267, include/cycle_searcher.h:45-53
// The recursive DFS passes are run with explicit stacks (same visit order, no stack-depth limit),
// the FVS acyclicity test is an incremental union-find (same accept/reject sequence, O(m a(n))
// instead of O(n m)), and the candidate filter is a top-down pass per tree (same accepted set as
// the per-edge LCA walk of include/Dijkstra.h:205-254).

#include "mcb_host.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <queue>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------------------------------------
// Graph containers
// ---------------------------------------------------------------------------------------------

// Simple input graph in CSR form, both directions stored (csr_graph, include/CsrGraph.h:12-218).
struct InputGraph {
    int32_t n = 0;
    std::vector<uint32_t> rows, cols, ro, deg;
    std::vector<int32_t> w;
};

// Reduced multigraph with the opposite-direction link of every edge (csr_multi_graph).
struct Multigraph {
    int32_t n = 0;
    std::vector<uint32_t> rows, cols, ro, deg, rev;
    std::vector<int32_t> w;
};

struct RowColLess {
    const uint32_t *r, *c;
    bool operator()(uint32_t a, uint32_t b) const {
        if (r[a] == r[b]) return c[a] < c[b];
        return r[a] < r[b];
    }
};

// Sorts edges by (row, col) with std::sort over the insertion sequence, then builds row offsets
// and degrees (include/CsrGraph.h:113-174).
void finish_input_csr(InputGraph &g) {
    const size_t m2 = g.rows.size();
    std::vector<uint32_t> idx(m2);
    std::iota(idx.begin(), idx.end(), 0u);
    std::sort(idx.begin(), idx.end(), RowColLess{g.rows.data(), g.cols.data()});
    std::vector<uint32_t> r(m2), c(m2);
    std::vector<int32_t> w(m2);
    for (size_t i = 0; i < m2; ++i) {
        r[i] = g.rows[idx[i]];
        c[i] = g.cols[idx[i]];
        w[i] = g.w[idx[i]];
    }
    g.rows.swap(r);
    g.cols.swap(c);
    g.w.swap(w);
    g.ro.assign(g.n + 1, 0);
    g.deg.assign(g.n, 0);
    for (size_t i = 0; i < m2; ++i) g.ro[g.rows[i]]++;
    uint32_t prev = 0;
    for (int32_t i = 0; i <= g.n; ++i) {
        uint32_t cur = g.ro[i];
        g.ro[i] = prev;
        prev += cur;
    }
    for (int32_t i = 0; i < g.n; ++i) g.deg[i] = g.ro[i + 1] - g.ro[i];
}

// Appends an undirected edge as two directed edges linked to each other
// (csr_multi_graph::insert, include/CsrGraphMulti.h:75-90).
void multi_insert(Multigraph &g, uint32_t a, uint32_t b, int32_t wt) {
    const uint32_t k = static_cast<uint32_t>(g.rows.size());
    g.rows.push_back(a);
    g.cols.push_back(b);
    g.w.push_back(wt);
    g.rev.push_back(k + 1);
    g.rows.push_back(b);
    g.cols.push_back(a);
    g.w.push_back(wt);
    g.rev.push_back(k);
}

// Sort by (row, col) with std::sort, carrying the opposite-direction links
// (include/CsrGraphMulti.h:96-168).
void finish_multi_csr(Multigraph &g) {
    const size_t m2 = g.rows.size();
    std::vector<uint32_t> idx(m2);
    std::iota(idx.begin(), idx.end(), 0u);
    std::sort(idx.begin(), idx.end(), RowColLess{g.rows.data(), g.cols.data()});
    std::vector<uint32_t> newpos(m2);
    for (size_t i = 0; i < m2; ++i) newpos[idx[i]] = static_cast<uint32_t>(i);
    std::vector<uint32_t> r(m2), c(m2), rv(m2);
    std::vector<int32_t> w(m2);
    for (size_t i = 0; i < m2; ++i) {
        r[i] = g.rows[idx[i]];
        c[i] = g.cols[idx[i]];
        w[i] = g.w[idx[i]];
        rv[i] = newpos[g.rev[idx[i]]];
    }
    g.rows.swap(r);
    g.cols.swap(c);
    g.w.swap(w);
    g.rev.swap(rv);
    g.ro.assign(g.n + 1, 0);
    g.deg.assign(g.n, 0);
    for (size_t i = 0; i < m2; ++i) g.ro[g.rows[i]]++;
    uint32_t prev = 0;
    for (int32_t i = 0; i <= g.n; ++i) {
        uint32_t cur = g.ro[i];
        g.ro[i] = prev;
        prev += cur;
    }
    for (int32_t i = 0; i < g.n; ++i) g.deg[i] = g.ro[i + 1] - g.ro[i];
}

// ---------------------------------------------------------------------------------------------
// Degree-two chain contraction (include/CsrGraphDegreeTwoRemoval.cpp:12-126)
// ---------------------------------------------------------------------------------------------

struct ChainEdge {
    uint32_t a, b;
    uint32_t weight;
};

// Depth-first pass from the first vertex of degree > 2.  An edge is marked for removal when it
// touches a degree-two vertex; a chain is closed when the walk leaves its last degree-two vertex.
// Post-order matters: the chain bookkeeping for the edge (row -> child) runs after the child's
// subtree is finished, exactly where the recursive reference resumes.
int mark_degree_two_chains(const InputGraph &g, std::vector<uint32_t> &removal,
                           std::vector<std::vector<uint32_t>> &chains, int32_t &src) {
    const int32_t n = g.n;
    src = -1;
    for (int32_t i = 0; i < n; ++i)
        if (g.deg[i] > 2) {
            src = i;
            break;
        }
    if (src < 0) return fail(MCB_E_GRAPH, "no vertex of degree > 2 to start the chain walk from");

    std::vector<uint8_t> visited(n, 0);
    std::vector<int64_t> parent(n, -1);
    bool have_temp = false;
    std::vector<uint32_t> temp;

    auto chain_step = [&](uint32_t row, uint32_t column, uint32_t offset) -> int {
        const bool dc = g.deg[column] == 2, dr = g.deg[row] == 2;
        if (dc && dr) {
            if (!have_temp) return fail(MCB_E_GRAPH, "degree-two chain walk entered a chain mid-way");
            temp.push_back(offset);
            removal.push_back(offset);
        } else if (dr) {
            temp.assign(1, offset);
            have_temp = true;
            removal.push_back(offset);
        } else if (dc) {
            if (!have_temp) return fail(MCB_E_GRAPH, "degree-two chain walk closed a chain it never opened");
            temp.push_back(offset);
            chains.push_back(temp);
            temp.clear();
            have_temp = false;
            removal.push_back(offset);
        }
        return MCB_OK;
    };

    struct Frame {
        uint32_t row, off;
        bool returning;
    };
    std::vector<Frame> stack;
    visited[src] = 1;
    stack.push_back({static_cast<uint32_t>(src), g.ro[src], false});
    while (!stack.empty()) {
        Frame &f = stack.back();
        const uint32_t row = f.row;
        if (f.returning) {
            f.returning = false;
            const uint32_t off = f.off++;
            int rc = chain_step(row, g.cols[off], off);
            if (rc) return rc;
            continue;
        }
        if (f.off >= g.ro[row + 1]) {
            stack.pop_back();
            continue;
        }
        const uint32_t off = f.off;
        const uint32_t column = g.cols[off];
        if (!visited[column]) {
            visited[column] = 1;
            parent[column] = row;
            f.returning = true;
            stack.push_back({column, g.ro[column], false});
            continue;
        }
        f.off++;
        if (static_cast<int64_t>(column) == parent[row]) {
            if (g.deg[column] == 2 || g.deg[row] == 2) removal.push_back(off);
            continue;
        }
        if (g.deg[column] == 2) {
            removal.push_back(off);
            continue;
        }
        int rc = chain_step(row, column, off);
        if (rc) return rc;
    }
    return MCB_OK;
}

// Relabel by first appearance and rebuild as a multigraph, new chain edges first
// (csr_multi_graph::get_modified_graph, include/CsrGraphMulti.h:229-300).
void build_reduced(const InputGraph &g, const std::vector<uint32_t> *removal,
                   const std::vector<ChainEdge> *new_edges, Multigraph &out) {
    const size_t m2 = g.rows.size();
    std::vector<uint8_t> filter(m2, 0);
    if (removal)
        for (uint32_t off : *removal) filter[off] = 1;
    std::vector<int64_t> label(g.n, -1);
    int64_t count = 0;
    for (size_t i = 0; i < m2; ++i) {
        if (filter[i]) continue;
        if (label[g.rows[i]] < 0) label[g.rows[i]] = count++;
        if (label[g.cols[i]] < 0) label[g.cols[i]] = count++;
    }
    if (new_edges)
        for (const ChainEdge &e : *new_edges) {
            if (label[e.a] < 0) label[e.a] = count++;
            if (label[e.b] < 0) label[e.b] = count++;
        }
    out = Multigraph();
    out.n = static_cast<int32_t>(count);
    if (new_edges)
        for (const ChainEdge &e : *new_edges)
            multi_insert(out, static_cast<uint32_t>(label[e.a]), static_cast<uint32_t>(label[e.b]),
                         static_cast<int32_t>(e.weight));
    for (size_t i = 0; i < m2; ++i)
        if (!filter[i] && g.rows[i] < g.cols[i])
            multi_insert(out, static_cast<uint32_t>(label[g.rows[i]]), static_cast<uint32_t>(label[g.cols[i]]),
                         g.w[i]);
    finish_multi_csr(out);
}

// ---------------------------------------------------------------------------------------------
// Spanning tree and cycle-space coordinates (include/CsrGraphMultiSpanningTree.cpp:11-135)
// ---------------------------------------------------------------------------------------------

int spanning_tree_coordinates(const Multigraph &g, int32_t src, std::vector<uint32_t> &nt_edges) {
    const int32_t n = g.n;
    if (src < 0 || src >= n) return fail(MCB_E_GRAPH, "spanning-tree root is not a vertex of the reduced graph");
    std::vector<uint8_t> visited(n, 0), is_tree(g.rows.size(), 0);
    std::vector<int64_t> parent(n, -1);
    int64_t tree_size = 0;
    struct Frame {
        uint32_t row, off;
    };
    std::vector<Frame> stack;
    visited[src] = 1;
    stack.push_back({static_cast<uint32_t>(src), g.ro[src]});
    while (!stack.empty()) {
        Frame &f = stack.back();
        const uint32_t row = f.row;
        if (f.off >= g.ro[row + 1]) {
            stack.pop_back();
            continue;
        }
        const uint32_t off = f.off++;
        const uint32_t column = g.cols[off];
        if (!visited[column]) {
            visited[column] = 1;
            ++tree_size;
            parent[column] = row;
            is_tree[off] = 1;
            stack.push_back({column, g.ro[column]});
            continue;
        }
        const uint32_t rv = g.rev[off];
        if (static_cast<int64_t>(column) == parent[row]) {
            if (is_tree[rv] == 1) {
                is_tree[off] = 1;
            } else if (is_tree[rv] == 2) {
                is_tree[off] = 2;
            } else {
                is_tree[off] = 2;
                is_tree[rv] = 2;
                nt_edges.push_back(row < column ? off : rv);
            }
        } else if (is_tree[off] == 0) {
            is_tree[off] = 2;
            is_tree[rv] = 2;
            nt_edges.push_back(row < column ? off : rv);
        }
    }
    if (tree_size != n - 1) return fail(MCB_E_GRAPH, "reduced graph is not connected");
    return MCB_OK;
}

// ---------------------------------------------------------------------------------------------
// Feedback vertex set (include/FVS.cpp)
// ---------------------------------------------------------------------------------------------

struct UnionFind {
    std::vector<int32_t> p;
    explicit UnionFind(int32_t n) : p(n) { std::iota(p.begin(), p.end(), 0); }
    int32_t find(int32_t x) {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
};

// The reference greedy (FVS::MGA) picks the vertex of least W/degree; its pruning step marks only
// the picked vertex dead (FVS.cpp:41-56), so W and degrees never change and the pick order is
// (degree desc, id asc).  The list is then reversed and each vertex is dropped from the set when
// the remaining graph stays acyclic (FVS.cpp:194-213).  Dropping v keeps the forest acyclic iff
// no two of v's edges into non-set vertices land in the same forest component (a parallel edge
// lands twice in one component); the forest only grows, so one union-find pass decides it.
int fvs_reverse_delete_impl(int32_t n, const uint32_t *ro, const uint32_t *cols, const int32_t *order,
                            int32_t *fvs_rank) {
    std::vector<int32_t> ord;
    if (order) {
        ord.assign(order, order + n);
        std::vector<uint8_t> seen(n, 0);
        for (int32_t v : ord) {
            if (v < 0 || v >= n || seen[v]) return fail(MCB_E_INVALID, "FVS order is not a permutation of the vertices");
            seen[v] = 1;
        }
    } else {
        ord.resize(n);
        std::iota(ord.begin(), ord.end(), 0);
        std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
            uint32_t da = ro[a + 1] - ro[a], db = ro[b + 1] - ro[b];
            if (da != db) return da < db;
            return a > b;
        });
    }
    std::vector<uint8_t> in_set(n, 1);
    UnionFind uf(n);
    std::vector<int32_t> comps;
    for (int32_t v : ord) {
        comps.clear();
        bool cycle = false;
        for (uint32_t e = ro[v]; e < ro[v + 1]; ++e) {
            uint32_t c = cols[e];
            if (c == static_cast<uint32_t>(v)) {  // a self-loop is a cycle on its own
                cycle = true;
                break;
            }
            if (in_set[c]) continue;
            comps.push_back(uf.find(static_cast<int32_t>(c)));
        }
        if (!cycle) {
            std::sort(comps.begin(), comps.end());
            cycle = std::adjacent_find(comps.begin(), comps.end()) != comps.end();
        }
        if (cycle) continue;  // v stays in the set
        in_set[v] = 0;
        for (int32_t r : comps) uf.p[uf.find(r)] = v;
    }
    int32_t count = 0;
    for (int32_t i = 0; i < n; ++i) fvs_rank[i] = in_set[i] ? count++ : -1;
    return count;
}

// ---------------------------------------------------------------------------------------------
// Per-root shortest-path trees and candidates
// ---------------------------------------------------------------------------------------------

struct HeapLess {  // include/Dijkstra.h:37-43: the heap's top is the least distance
    bool operator()(const std::pair<int, int> &a, const std::pair<int, int> &b) const {
        return a.second > b.second;
    }
};

// The binary heap of the search, as std::priority_queue keeps it (a max-heap under the comparator
// "larger distance is less", so the top holds the least distance), with its moves written out so
// that the child choice is branch-free.  The element moves are those of std::push_heap /
// std::pop_heap in libstdc++, which fix the pop order among equal distances: push sifts the new
// element up while its parent is strictly less (larger distance); pop moves the last element out,
// walks the hole from the top to a leaf always into the greater child (the right one unless it is
// strictly less than the left; a lone last child at an even length), and sifts the moved element up
// from there.  The full-tree searches keep std::priority_queue itself, and the tests require both to
// give the same trees.
struct SearchHeap {
    std::vector<std::pair<int, int>> a;  // (vertex, distance)

    void clear() { a.clear(); }
    bool empty() const { return a.empty(); }
    int top_vertex() const { return a[0].first; }
    static bool less(const std::pair<int, int> &x, const std::pair<int, int> &y) { return x.second > y.second; }

    void sift_up(size_t hole, size_t top, std::pair<int, int> v) {
        while (hole > top) {
            const size_t parent = (hole - 1) / 2;
            if (!less(a[parent], v)) break;
            a[hole] = a[parent];
            hole = parent;
        }
        a[hole] = v;
    }
    void push(int vertex, int dist) {
        a.emplace_back(vertex, dist);
        sift_up(a.size() - 1, 0, a.back());
    }
    void pop() {
        const size_t len = a.size() - 1;  // the heap after removing the top
        if (len == 0) {
            a.pop_back();
            return;
        }
        const std::pair<int, int> v = a[len];
        a[len] = a[0];  // the old top goes to the end, then off
        size_t hole = 0, child = 0;
        if (len > 2) {
            const size_t lim = (len - 1) / 2;
            while (child < lim) {
                child = 2 * (child + 1);
                child -= static_cast<size_t>(a[child].second > a[child - 1].second);
                a[hole] = a[child];
                hole = child;
            }
        }
        if ((len & 1) == 0 && child == (len - 2) / 2) {
            child = 2 * (child + 1);
            a[hole] = a[child - 1];
            hole = child - 1;
        }
        sift_up(hole, 0, v);
        a.pop_back();
    }
};

struct StdHeap {
    std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>, HeapLess> q;
    void clear() { q = decltype(q)(); }
    bool empty() const { return q.empty(); }
    int top_vertex() const { return q.top().first; }
    void push(int vertex, int dist) { q.push(std::make_pair(vertex, dist)); }
    void pop() { q.pop(); }
};

// One vertex of a shortest-path search, kept together because the relax loop reads them together.
struct VertexState {
    int32_t dist;    // -1: not reached
    int32_t pedge;   // offset of the edge parent -> vertex (-1 for the root and unreached vertices)
    int32_t level;   // hop depth; set when the vertex is finalised
    uint32_t branch; // the child of the root whose subtree holds the vertex (~0 for the root)
    uint8_t in_tree, blocked, tent_blocked, is_root;  // is_root: in the FVS (fixed per problem)
};

struct SpTree {
    std::vector<VertexState> v;
    std::vector<uint32_t> order;  // vertices in finalisation order, root excluded
};

// Per-thread scratch of the root loop.  Every per-vertex entry holds its idle value (dist, pedge,
// level -1; flags 0; branch, pos_of ~0) except at the vertices listed in `touched`, which the next
// root resets, so a root costs time in the part of the graph its search reached.
struct Workspace {
    SpTree t;
    std::vector<uint8_t> keep;
    std::vector<uint32_t> pos_of, touched, open_rows, level_count, level_start;
    SearchHeap heap;
    StdHeap std_heap;
    int32_t n = -1;

    void prepare(int32_t nodes, const int32_t *fvs_rank) {
        if (n != nodes) {
            n = nodes;
            t.v.assign(n, VertexState{-1, -1, -1, 0xFFFFFFFFu, 0, 0, 0, 0});
            for (int32_t i = 0; i < n; ++i) t.v[i].is_root = fvs_rank[i] >= 0 ? 1 : 0;
            keep.assign(n, 0);
            pos_of.assign(n, 0xFFFFFFFFu);
        } else {
            for (uint32_t x : touched) {
                VertexState &s = t.v[x];
                s.dist = s.pedge = s.level = -1;
                s.branch = 0xFFFFFFFFu;
                s.in_tree = s.blocked = s.tent_blocked = 0;
                keep[x] = 0;
                pos_of[x] = 0xFFFFFFFFu;
            }
        }
        touched.clear();
        t.order.clear();
    }
};

// dijkstra::dijkstra_sp (include/Dijkstra.h:78-127): lazy-deletion binary heap, strict '<' relax.
// Heap: std::priority_queue's (StdHeap) or SearchHeap, which pops in the same order.  At each
// pop the vertex's parent is final, so blocked (the tree path from src holds a root of smaller id
// than src: include/Dijkstra.h:205-254 rejects every cycle through such a vertex), branch and level
// are set there.
//
// stop_when_closed: stop as soon as every reached, unfinalised vertex is blocked through its current
// parent.  From then on every vertex finalised is blocked (its final parent is finalised before it,
// and by induction blocked), so no later vertex can be a candidate endpoint or lie on the path of
// one.  The pops up to that point are the full search's pops (same heap, same sequence), so the
// finalised part of the tree, its order and its distances are the full search's.
template <typename Heap>
void dijkstra(const Multigraph &g, uint32_t src, bool stop_when_closed, Workspace &ws, Heap &heap) {
    VertexState *V = ws.t.v.data();
    const uint32_t *ro = g.ro.data(), *cols = g.cols.data(), *rows = g.rows.data();
    const int32_t *wt = g.w.data();
    V[src].dist = 0;
    V[src].level = 0;
    ws.touched.push_back(src);
    int64_t open = 1;  // reached, unfinalised vertices whose current path is not blocked
    heap.clear();
    heap.push(static_cast<int>(src), 0);
    while (!heap.empty()) {
        const uint32_t u = static_cast<uint32_t>(heap.top_vertex());
        heap.pop();
        VertexState &su = V[u];
        if (su.in_tree) continue;
        su.in_tree = 1;
        if (u != src) {
            ws.t.order.push_back(u);
            const uint32_t p = rows[su.pedge];
            const VertexState &sp = V[p];
            su.blocked = (sp.blocked || (su.is_root && u < src)) ? 1 : 0;
            su.branch = (p == src) ? u : sp.branch;
            su.level = sp.level + 1;
        }
        const bool bu = su.blocked != 0;
        if (!bu) --open;
        const int32_t du = su.dist;
        for (uint32_t off = ro[u]; off < ro[u + 1]; ++off) {
            const uint32_t c = cols[off];
            VertexState &sc = V[c];
            if (sc.in_tree) continue;
            const int32_t cand = du + wt[off];
            if (sc.dist == -1 || cand < sc.dist) {
                const uint8_t nb = (bu || (sc.is_root && c < src)) ? 1 : 0;
                if (sc.dist == -1) {
                    ws.touched.push_back(c);
                    open += !nb;
                } else {
                    open += static_cast<int64_t>(!nb) - static_cast<int64_t>(!sc.tent_blocked);
                }
                sc.tent_blocked = nb;
                sc.dist = cand;
                sc.pedge = static_cast<int32_t>(off);
                heap.push(static_cast<int>(c), cand);
            }
        }
        if (stop_when_closed && open == 0) break;
    }
}

struct RootOutput {
    // candidates in the order the reference collects them for this root (map iteration order)
    std::vector<uint32_t> cand_edge;
    std::vector<int32_t> cand_weight;
    std::vector<uint32_t> cand_pu, cand_pv;
    // position layout
    std::vector<uint32_t> parent_pos;
    std::vector<int32_t> coord;
    std::vector<uint32_t> vertex;  // only with keep_vertex_trees
    int32_t depth = 0;
};


}  // namespace

struct mcb_problem {
    mcb_problem_info info{};
    Multigraph g;
    std::vector<int32_t> fvs_rank;
    std::vector<uint32_t> roots, nt_edges;
    std::vector<int32_t> nt_map;
    std::vector<uint32_t> cand_root, cand_edge, cand_pu, cand_pv;
    std::vector<int32_t> cand_weight, cand_coord;
    std::vector<int64_t> tree_offset, root_cand_offset;
    std::vector<uint32_t> parent_pos, root_cand_rank;
    std::vector<int32_t> pos_coord, vertex_parent;
    std::vector<uint32_t> pos_vertex;
    bool trees_pending = false;  // built with external_trees: mcb_problem_finish_trees completes it
};

namespace {

// One root: Dijkstra, SP non-tree edges (include/Dijkstra.h:129-149), the candidate filter of
// include/Dijkstra.h:205-254 and the grouping of include/cycle_searcher.h:45-53.  The pruned layout
// without per-vertex parent rows stops the search once the rest of the graph is blocked (see
// dijkstra); the unpruned layout and the parent rows need the whole tree.
void process_root(const mcb_problem &P, uint32_t src, bool prune, Workspace &ws, RootOutput &out,
                  int32_t *vertex_parent_row) {
    const Multigraph &g = P.g;
    const int32_t n = g.n;
    SpTree &t = ws.t;
    ws.prepare(n, P.fvs_rank.data());
    if (prune && !vertex_parent_row)
        dijkstra(g, src, true, ws, ws.heap);
    else
        dijkstra(g, src, false, ws, ws.std_heap);
    const VertexState *V = t.v.data();
    if (vertex_parent_row)
        for (int32_t v = 0; v < n; ++v) vertex_parent_row[v] = V[v].pedge;

    // Accept (u,v): the tree paths from u and v meet first at src and avoid smaller roots.  The
    // reference lists the SP non-tree edges in offset order, each at the first of its two directions
    // (include/Dijkstra.h:135-148), and filters that list.  An accepted edge has both ends finalised
    // and unblocked, and its first direction lies in the row of its smaller end, so scanning the rows
    // of the finalised unblocked vertices in ascending id, and in each row the edges to a larger (or,
    // for a loop, the same) vertex, meets the accepted edges in the same order.  An edge is a tree
    // edge when it is the parent edge of either end.
    ws.open_rows.clear();
    ws.open_rows.push_back(src);
    for (uint32_t v : t.order)
        if (!V[v].blocked) ws.open_rows.push_back(v);
    if (ws.open_rows.size() * 16 < static_cast<size_t>(n)) {
        std::sort(ws.open_rows.begin(), ws.open_rows.end());
    } else {  // most of the graph: list the finalised unblocked vertices by a pass over the ids
        ws.open_rows.clear();
        for (int32_t v = 0; v < n; ++v)
            if (V[v].in_tree && !V[v].blocked) ws.open_rows.push_back(static_cast<uint32_t>(v));
    }
    std::unordered_map<unsigned long long, uint32_t> groups_of;
    std::vector<std::vector<uint32_t>> groups;
    for (uint32_t u : ws.open_rows) {
        for (uint32_t e = g.ro[u]; e < g.ro[u + 1]; ++e) {
            const uint32_t v = g.cols[e];
            if (v < u || (v == u && e > g.rev[e])) continue;  // met from the other direction
            if (!V[v].in_tree || V[v].blocked) continue;
            if (V[v].pedge == static_cast<int32_t>(e) || V[u].pedge == static_cast<int32_t>(g.rev[e])) continue;
            if (!(u == src || v == src || V[u].branch != V[v].branch)) continue;
            const unsigned long long key = (static_cast<unsigned long long>(u) << 32) | v;
            auto it = groups_of.find(key);
            if (it == groups_of.end()) {
                groups_of.insert(std::make_pair(key, static_cast<uint32_t>(groups.size())));
                groups.emplace_back(1, e);
            } else {
                groups[it->second].push_back(e);
            }
        }
    }
    out.cand_edge.clear();
    out.cand_weight.clear();
    for (auto &kv : groups_of)
        for (uint32_t e : groups[kv.second]) {
            out.cand_edge.push_back(e);
            out.cand_weight.push_back(V[g.rows[e]].dist + V[g.cols[e]].dist + g.w[e]);
        }

    // Which vertices get a position: all of them, or the closure of the candidate endpoints
    // under the parent relation (every ancestor of a kept vertex is kept).
    if (!prune)
        for (int32_t v = 0; v < n; ++v) ws.keep[v] = 1;
    ws.keep[src] = 1;
    if (prune) {
        for (uint32_t e : out.cand_edge) {
            for (uint32_t x : {g.rows[e], g.cols[e]}) {
                while (!ws.keep[x]) {
                    ws.keep[x] = 1;
                    x = g.rows[V[x].pedge];
                }
            }
        }
    }
    // BFS-level layout: level buckets, finalisation order inside a level (parents come first).
    int32_t maxlev = 0;
    for (uint32_t v : t.order)
        if (ws.keep[v]) maxlev = std::max(maxlev, V[v].level);
    ws.level_count.assign(maxlev + 2, 0);
    ws.level_count[0] = 1;
    for (uint32_t v : t.order)
        if (ws.keep[v]) ws.level_count[V[v].level]++;
    ws.level_start.assign(maxlev + 2, 0);
    for (int32_t l = 1; l <= maxlev + 1; ++l) ws.level_start[l] = ws.level_start[l - 1] + ws.level_count[l - 1];
    const uint32_t npos = ws.level_start[maxlev + 1];
    ws.pos_of[src] = 0;
    ws.level_start[0] = 1;
    for (uint32_t v : t.order)
        if (ws.keep[v]) ws.pos_of[v] = ws.level_start[V[v].level]++;
    out.parent_pos.assign(npos, 0xFFFFFFFFu);
    out.coord.assign(npos, -1);
    for (uint32_t v : t.order) {
        if (!ws.keep[v]) continue;
        const uint32_t pos = ws.pos_of[v];
        const uint32_t e = static_cast<uint32_t>(V[v].pedge);
        out.parent_pos[pos] = ws.pos_of[g.rows[e]];
        out.coord[pos] = P.nt_map[e];
    }
    out.depth = maxlev;
    if (vertex_parent_row) {
        out.vertex.assign(npos, 0);
        out.vertex[0] = src;
        for (uint32_t v : t.order)
            if (ws.keep[v]) out.vertex[ws.pos_of[v]] = v;
    }
    out.cand_pu.resize(out.cand_edge.size());
    out.cand_pv.resize(out.cand_edge.size());
    for (size_t k = 0; k < out.cand_edge.size(); ++k) {
        out.cand_pu[k] = ws.pos_of[g.rows[out.cand_edge[k]]];
        out.cand_pv[k] = ws.pos_of[g.cols[out.cand_edge[k]]];
    }
}

// One root's outputs from the results of its stopped search, as a device search hands them over: the
// accepted candidate edges with their weights and the vertices of the pruned layout other than the root
// (vertex, parent edge, level, finalisation index), both in any order.  The accepted edges are listed in
// ascending offset, which is the order process_root meets them (rows of ascending vertex, each row in
// offset order), so the groups and their iteration order are process_root's; the layout puts the kept
// vertices in level buckets, finalisation order inside a level, as process_root does.
void root_from_search(const mcb_problem &P, uint32_t src, int32_t nc, const uint32_t *ce, const int32_t *cw,
                      int32_t nk, const uint32_t *kept, Workspace &ws, RootOutput &out) {
    const Multigraph &g = P.g;
    ws.prepare(g.n, P.fvs_rank.data());
    std::vector<std::pair<uint32_t, int32_t>> cands(nc);
    for (int32_t i = 0; i < nc; ++i) cands[i] = std::make_pair(ce[i], cw[i]);
    std::sort(cands.begin(), cands.end());
    std::unordered_map<unsigned long long, uint32_t> groups_of;
    std::vector<std::vector<uint32_t>> groups;  // indices into cands
    for (uint32_t i = 0; i < cands.size(); ++i) {
        const uint32_t e = cands[i].first;
        const unsigned long long key = (static_cast<unsigned long long>(g.rows[e]) << 32) | g.cols[e];
        auto it = groups_of.find(key);
        if (it == groups_of.end()) {
            groups_of.insert(std::make_pair(key, static_cast<uint32_t>(groups.size())));
            groups.emplace_back(1, i);
        } else {
            groups[it->second].push_back(i);
        }
    }
    out.cand_edge.clear();
    out.cand_weight.clear();
    for (auto &kv : groups_of)
        for (uint32_t i : groups[kv.second]) {
            out.cand_edge.push_back(cands[i].first);
            out.cand_weight.push_back(cands[i].second);
        }
    // (level, finalisation index, vertex, parent edge), sorted: level buckets, finalisation order inside
    struct Kept {
        uint32_t level, ord, v, pedge;
        bool operator<(const Kept &o) const { return level != o.level ? level < o.level : ord < o.ord; }
    };
    std::vector<Kept> ks(nk);
    for (int32_t i = 0; i < nk; ++i) ks[i] = Kept{kept[4 * i + 2], kept[4 * i + 3], kept[4 * i], kept[4 * i + 1]};
    std::sort(ks.begin(), ks.end());
    const uint32_t npos = static_cast<uint32_t>(nk) + 1;
    ws.pos_of[src] = 0;
    ws.touched.push_back(src);
    for (uint32_t i = 0; i < ks.size(); ++i) {
        ws.pos_of[ks[i].v] = i + 1;
        ws.touched.push_back(ks[i].v);
    }
    out.parent_pos.assign(npos, 0xFFFFFFFFu);
    out.coord.assign(npos, -1);
    for (uint32_t i = 0; i < ks.size(); ++i) {
        out.parent_pos[i + 1] = ws.pos_of[g.rows[ks[i].pedge]];
        out.coord[i + 1] = P.nt_map[ks[i].pedge];
    }
    out.depth = ks.empty() ? 0 : static_cast<int32_t>(ks.back().level);
    out.cand_pu.resize(out.cand_edge.size());
    out.cand_pv.resize(out.cand_edge.size());
    for (size_t k = 0; k < out.cand_edge.size(); ++k) {
        out.cand_pu[k] = ws.pos_of[g.rows[out.cand_edge[k]]];
        out.cand_pv[k] = ws.pos_of[g.cols[out.cand_edge[k]]];
    }
}

// The host's stopped search of one root in the form root_from_search takes (the finalisation index
// counts the root as 0).
void search_results(const mcb_problem &P, uint32_t src, Workspace &ws, std::vector<uint32_t> &ce,
                    std::vector<int32_t> &cw, std::vector<uint32_t> &kept) {
    const Multigraph &g = P.g;
    ws.prepare(g.n, P.fvs_rank.data());
    dijkstra(g, src, true, ws, ws.heap);
    const VertexState *V = ws.t.v.data();
    ce.clear();
    cw.clear();
    kept.clear();
    for (uint32_t i = 0; i < ws.t.order.size(); ++i) ws.pos_of[ws.t.order[i]] = i + 1;  // finalisation index
    std::vector<uint32_t> open_rows(1, src);
    for (uint32_t v : ws.t.order)
        if (!V[v].blocked) open_rows.push_back(v);
    std::sort(open_rows.begin(), open_rows.end());
    for (uint32_t u : open_rows)
        for (uint32_t e = g.ro[u]; e < g.ro[u + 1]; ++e) {
            const uint32_t v = g.cols[e];
            if (v < u || (v == u && e > g.rev[e])) continue;
            if (!V[v].in_tree || V[v].blocked) continue;
            if (V[v].pedge == static_cast<int32_t>(e) || V[u].pedge == static_cast<int32_t>(g.rev[e])) continue;
            if (!(u == src || v == src || V[u].branch != V[v].branch)) continue;
            ce.push_back(e);
            cw.push_back(V[u].dist + V[v].dist + g.w[e]);
        }
    ws.keep[src] = 1;
    for (uint32_t e : ce)
        for (uint32_t x : {g.rows[e], g.cols[e]})
            while (!ws.keep[x]) {
                ws.keep[x] = 1;
                kept.insert(kept.end(), {x, static_cast<uint32_t>(V[x].pedge), static_cast<uint32_t>(V[x].level),
                                         ws.pos_of[x]});
                x = g.rows[V[x].pedge];
            }
}

// Stage times of build_from_input on stderr when MCB_HOST_TIMING is set (seconds since the previous mark).
struct StageClock {
    bool on = std::getenv("MCB_HOST_TIMING") != nullptr;
    double t = omp_get_wtime();
    void mark(const char *stage) {
        if (!on) return;
        const double now = omp_get_wtime();
        std::fprintf(stderr, "[mcb_host] %-28s %8.3f s\n", stage, now - t);
        t = now;
    }
};

int assemble(mcb_problem &P, std::vector<RootOutput> &per_root, bool keep_vertex_trees, int threads,
             StageClock &clock);

int build_from_input(InputGraph &in, int64_t input_weight_sum, const mcb_build_options &opt, mcb_problem **out) {
    StageClock clock;
    std::unique_ptr<mcb_problem> P(new (std::nothrow) mcb_problem());
    if (!P) return fail(MCB_E_NOMEM, "out of memory");
    finish_input_csr(in);
    mcb_problem_info &info = P->info;
    info.input_nodes = in.n;
    info.input_edges = static_cast<int32_t>(in.rows.size() / 2);
    info.input_total_weight = input_weight_sum;
    for (uint32_t i = 0; i < in.rows.size(); ++i)
        if (in.rows[i] == in.cols[i]) return fail(MCB_E_GRAPH, "input graph has a self-loop");

    int32_t deg2 = 0;
    for (int32_t i = 0; i < in.n; ++i) deg2 += (in.deg[i] == 2);
    if (deg2 == in.n) {  // the whole graph is one cycle (mcb_cycle_deg_two_remove.cpp:104-113)
        info.single_cycle = 1;
        info.nodes_removed = in.n - 1;
        info.num_non_tree = 1;
        info.words = 1;
        info.fvs_size = 1;
        info.num_candidates = 1;
        info.nodes = in.n;
        info.directed_edges = static_cast<int32_t>(in.rows.size());
        *out = P.release();
        return MCB_OK;
    }

    int32_t src = 0;
    if (opt.reduce_degree_two) {
        std::vector<uint32_t> removal;
        std::vector<std::vector<uint32_t>> chains;
        int rc = mark_degree_two_chains(in, removal, chains, src);
        if (rc) return rc;
        std::vector<ChainEdge> new_edges;
        int32_t removed = 0;
        for (const auto &ch : chains) {
            uint32_t total = 0;
            for (uint32_t off : ch) total += static_cast<uint32_t>(in.w[off]);
            new_edges.push_back({in.rows[ch.back()], in.cols[ch.front()], total});
            removed += static_cast<int32_t>(ch.size()) - 1;
        }
        info.nodes_removed = removed;
        build_reduced(in, &removal, &new_edges, P->g);
    } else {
        build_reduced(in, nullptr, nullptr, P->g);
    }
    in = InputGraph();  // release the input graph early
    clock.mark("csr + degree-two contraction");

    Multigraph &g = P->g;
    info.source_vertex = src;
    info.nodes = g.n;
    info.directed_edges = static_cast<int32_t>(g.rows.size());

    // FVS and root numbering by vertex id.
    P->fvs_rank.assign(g.n, -1);
    const int32_t *order = nullptr;
    if (opt.fvs_order) {
        if (opt.fvs_order_len != g.n) return fail(MCB_E_INVALID, "FVS order length differs from the reduced vertex count");
        order = opt.fvs_order;
    }
    int F = fvs_reverse_delete_impl(g.n, g.ro.data(), g.cols.data(), order, P->fvs_rank.data());
    if (F < 0) return F;
    info.fvs_size = F;
    clock.mark("fvs");
    P->roots.resize(F);
    for (int32_t v = 0; v < g.n; ++v)
        if (P->fvs_rank[v] >= 0) P->roots[P->fvs_rank[v]] = static_cast<uint32_t>(v);
    if (opt.stop_after_fvs) {  // the reduced graph and its FVS only
        info.num_non_tree = static_cast<int32_t>(g.rows.size() / 2) - g.n + 1;
        info.words = (info.num_non_tree + 63) / 64;
        *out = P.release();
        return MCB_OK;
    }

    // Coordinates.
    int rc = spanning_tree_coordinates(g, src, P->nt_edges);
    if (rc) return rc;
    const int32_t N = static_cast<int32_t>(P->nt_edges.size());
    info.num_non_tree = N;
    info.words = (N + 63) / 64;
    P->nt_map.assign(g.rows.size(), -1);
    for (int32_t i = 0; i < N; ++i) P->nt_map[P->nt_edges[i]] = i;
    for (size_t i = 0; i < g.rows.size(); ++i)
        if (P->nt_map[i] < 0 && P->nt_map[g.rev[i]] >= 0) P->nt_map[i] = P->nt_map[g.rev[i]];

    clock.mark("spanning tree, coordinates");
    if (opt.external_trees) {
        if (!opt.prune_trees || opt.keep_vertex_trees)
            return fail(MCB_E_INVALID, "external trees need prune_trees=1 and keep_vertex_trees=0");
        P->trees_pending = true;
        *out = P.release();
        return MCB_OK;
    }
    // Per-root trees and candidates.
    std::vector<RootOutput> per_root(F);
    if (opt.keep_vertex_trees) P->vertex_parent.assign(static_cast<size_t>(F) * g.n, -1);
    int threads = opt.num_threads > 0 ? opt.num_threads : omp_get_max_threads();
    int err = 0;
#pragma omp parallel num_threads(threads)
    {
        Workspace ws;
#pragma omp for schedule(dynamic, 4)
        for (int32_t r = 0; r < F; ++r) {
            try {
                process_root(*P, P->roots[r], opt.prune_trees != 0, ws, per_root[r],
                             opt.keep_vertex_trees ? P->vertex_parent.data() + static_cast<size_t>(r) * g.n : nullptr);
            } catch (const std::bad_alloc &) {
#pragma omp atomic write
                err = 1;
            }
        }
    }
    if (err) return fail(MCB_E_NOMEM, "out of memory while building shortest-path trees");
    clock.mark("shortest-path trees");
    int rc2 = assemble(*P, per_root, opt.keep_vertex_trees != 0, threads, clock);
    if (rc2) return rc2;
    *out = P.release();
    return MCB_OK;
}

// Candidates of every root in the reference's collection order, ranked by weight, and the
// concatenated position layouts (mcb/mcb_cycle_deg_two_remove.cpp:244-267).
int assemble(mcb_problem &Pr, std::vector<RootOutput> &per_root, bool keep_vertex_trees, int threads,
             StageClock &clock) {
    mcb_problem *P = &Pr;
    mcb_problem_info &info = P->info;
    const int32_t F = info.fvs_size;
    // Collect candidates root by root (roots ascend by vertex id) and rank them by weight with
    // the same std::sort over the same sequence (mcb_cycle_deg_two_remove.cpp:247-258).  Everything
    // but the sort runs on the library's threads: each root's slice of the collection and of the
    // layout starts at a prefix sum, and the per-root rank lists are a stable counting sort.
    std::vector<int64_t> cand_at(static_cast<size_t>(F) + 1, 0);
    P->tree_offset.assign(static_cast<size_t>(F) + 1, 0);
    int32_t max_pos = 0, max_depth = 0;
    for (int32_t r = 0; r < F; ++r) {
        const RootOutput &o = per_root[r];
        cand_at[r + 1] = cand_at[r] + static_cast<int64_t>(o.cand_edge.size());
        P->tree_offset[r + 1] = P->tree_offset[r] + static_cast<int64_t>(o.parent_pos.size());
        max_pos = std::max<int32_t>(max_pos, static_cast<int32_t>(o.parent_pos.size()));
        max_depth = std::max(max_depth, o.depth);
    }
    const int64_t K = cand_at[F], positions = P->tree_offset[F];
    if (K > 0xFFFFFFFFll - 1) return fail(MCB_E_GRAPH, "more than 2^32-1 candidate cycles");
    info.num_candidates = K;
    info.positions = positions;
    info.max_positions = max_pos;
    info.max_depth = max_depth;

    // The sort runs on (weight, collection index) pairs compared by weight alone: std::sort's
    // control flow depends only on the outcomes of its comparisons, which are those of sorting the
    // indices by weight, so the permutation (ties included) is the same, without a gather per
    // comparison.
    struct Keyed {
        int32_t w;
        uint32_t k;
    };
    std::vector<Keyed> keyed(K);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (int32_t r = 0; r < F; ++r) {
        const RootOutput &o = per_root[r];
        for (size_t i = 0; i < o.cand_edge.size(); ++i) {
            const int64_t k = cand_at[r] + static_cast<int64_t>(i);
            keyed[k] = Keyed{o.cand_weight[i], static_cast<uint32_t>(k)};
        }
    }
    std::sort(keyed.begin(), keyed.end(), [](const Keyed &a, const Keyed &b) { return a.w < b.w; });
    clock.mark("candidate collect + sort");
    // collection index -> root, then the ranked arrays gathered from the roots' own lists
    std::vector<uint32_t> root_of(K);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (int32_t r = 0; r < F; ++r)
        std::fill(root_of.begin() + cand_at[r], root_of.begin() + cand_at[r + 1], static_cast<uint32_t>(r));
    P->cand_root.resize(K);
    P->cand_edge.resize(K);
    P->cand_weight.resize(K);
    P->cand_pu.resize(K);
    P->cand_pv.resize(K);
    P->cand_coord.resize(K);
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int64_t k = 0; k < K; ++k) {
        const uint32_t s = keyed[k].k, r = root_of[s];
        const RootOutput &o = per_root[r];
        const size_t i = static_cast<size_t>(s - cand_at[r]);
        P->cand_root[k] = r;
        P->cand_edge[k] = o.cand_edge[i];
        P->cand_weight[k] = o.cand_weight[i];
        P->cand_pu[k] = o.cand_pu[i];
        P->cand_pv[k] = o.cand_pv[i];
        P->cand_coord[k] = P->nt_map[o.cand_edge[i]];
    }
    std::vector<Keyed>().swap(keyed);
    std::vector<uint32_t>().swap(root_of);
    // Per-root candidate lists, ranks ascending: each root's list has exactly its own candidates, so
    // root r's slice starts at cand_at[r]; thread t's chunk of ranks starts inside each slice after the
    // earlier chunks' counts for that root.
    P->root_cand_offset.assign(cand_at.begin(), cand_at.end());
    P->root_cand_rank.resize(K);
    {
        const int T = std::max(1, std::min<int>(threads, static_cast<int>(K / 65536) + 1));
        std::vector<int64_t> cnt(static_cast<size_t>(T) * F, 0);
#pragma omp parallel num_threads(T)
        {
            const int t = omp_get_thread_num(), nt = omp_get_num_threads();  // nt <= T
            const int64_t chunk = (K + nt - 1) / nt;
            const int64_t k0 = std::min(K, chunk * t), k1 = std::min(K, k0 + chunk);
            int64_t *c = cnt.data() + static_cast<size_t>(t) * F;
            for (int64_t k = k0; k < k1; ++k) ++c[P->cand_root[k]];
#pragma omp barrier
#pragma omp for schedule(static)
            for (int32_t r = 0; r < F; ++r) {
                int64_t at = cand_at[r];
                for (int u = 0; u < omp_get_num_threads(); ++u) {
                    const int64_t x = cnt[static_cast<size_t>(u) * F + r];
                    cnt[static_cast<size_t>(u) * F + r] = at;
                    at += x;
                }
            }
            for (int64_t k = k0; k < k1; ++k) P->root_cand_rank[c[P->cand_root[k]]++] = static_cast<uint32_t>(k);
        }
    }
    clock.mark("rank arrays");
    // Concatenated position layouts.
    P->parent_pos.resize(positions);
    P->pos_coord.resize(positions);
    if (keep_vertex_trees) P->pos_vertex.resize(positions);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (int32_t r = 0; r < F; ++r) {
        RootOutput &o = per_root[r];
        std::copy(o.parent_pos.begin(), o.parent_pos.end(), P->parent_pos.begin() + P->tree_offset[r]);
        std::copy(o.coord.begin(), o.coord.end(), P->pos_coord.begin() + P->tree_offset[r]);
        if (keep_vertex_trees)
            std::copy(o.vertex.begin(), o.vertex.end(), P->pos_vertex.begin() + P->tree_offset[r]);
        o = RootOutput();
    }
    clock.mark("position layout");
    return MCB_OK;
}

bool read_line(FILE *f, std::string &line) {
    line.clear();
    int ch;
    while ((ch = fgetc(f)) != EOF) {
        if (ch == '\n') return true;
        line.push_back(static_cast<char>(ch));
    }
    return !line.empty();
}

std::string lower(std::string s) {
    for (char &c : s) c = static_cast<char>(tolower(static_cast<unsigned char>(c)));
    return s;
}

}  // namespace

extern "C" {

void mcb_build_options_init(mcb_build_options *opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->reduce_degree_two = 1;
    opt->prune_trees = 0;
    opt->keep_vertex_trees = 0;
    opt->num_threads = 0;
}

const char *mcb_host_last_error(void) { return g_err.c_str(); }

namespace {
// Features of the vertices of a CSR multigraph restricted to alive[v] != 0 (alive NULL: all).  With ids
// NULL, rows for every vertex (zeros for the others); else the rows of the alive vertices only, in
// ascending vertex order, with their ids.  Weight sums are sums of integers, exact in double, so the
// rows do not depend on which vertices the CSR leaves out (a compacted graph gives the same rows).
struct FeatureScratch {
    std::vector<int64_t> deg, two, d, core;
    std::vector<double> wsum;
    std::vector<int32_t> bin, pos, vert;
};

void vertex_features_csr(int32_t n, const uint32_t *ro, const uint32_t *cols, const int32_t *weights,
                         const uint8_t *alive, float *out, int32_t *ids, const int32_t *id_of, FeatureScratch &fs) {
    auto live = [&](uint32_t x) { return !alive || alive[x]; };
    std::vector<int64_t> &deg = fs.deg, &two = fs.two, &d = fs.d, &core = fs.core;
    std::vector<double> &wsum = fs.wsum;
    deg.assign(n, 0);
    two.assign(n, 0);
    wsum.assign(n, 0.0);
    // Per-vertex passes on the library's threads above a size where that pays; the weight total is
    // summed in vertex order afterwards (sums of integers, exact either way).
    const bool par = n >= (1 << 15);
#pragma omp parallel for schedule(static) if (par)
    for (int32_t v = 0; v < n; ++v) {
        if (!live(v)) continue;
        int64_t dv = 0;
        double wv = 0.0;
        for (uint32_t e = ro[v]; e < ro[v + 1]; ++e)
            if (live(cols[e])) {
                ++dv;
                wv += weights[e];
            }
        deg[v] = dv;
        wsum[v] = wv;
    }
    int64_t maxd = 0;
#pragma omp parallel for schedule(static) reduction(max : maxd) if (par)
    for (int32_t v = 0; v < n; ++v) {
        maxd = std::max(maxd, deg[v]);
        if (!live(v)) continue;
        int64_t t = 0;
        for (uint32_t e = ro[v]; e < ro[v + 1]; ++e)
            if (live(cols[e])) t += deg[cols[e]];
        two[v] = t;
    }
    double wall = 0.0;
    int64_t ends = 0;
    for (int32_t v = 0; v < n; ++v) {
        wall += wsum[v];
        ends += deg[v];
    }
    const double wmean = ends ? wall / static_cast<double>(ends) : 1.0;
    // core numbers: peel vertices in order of current degree (bucket queue; parallel edges count)
    d.assign(deg.begin(), deg.end());
    core.assign(n, 0);
    std::vector<int32_t> &bin = fs.bin, &pos = fs.pos, &vert = fs.vert;
    bin.assign(static_cast<size_t>(maxd) + 2, 0);
    pos.resize(n);
    vert.resize(n);
    for (int32_t v = 0; v < n; ++v) ++bin[d[v]];
    int32_t start = 0;
    for (int64_t k = 0; k <= maxd; ++k) {
        const int32_t c = bin[k];
        bin[k] = start;
        start += c;
    }
    for (int32_t v = 0; v < n; ++v) {
        pos[v] = bin[d[v]]++;
        vert[pos[v]] = v;
    }
    for (int64_t k = maxd; k > 0; --k) bin[k] = bin[k - 1];
    bin[0] = 0;
    int64_t maxcore = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t v = vert[i];
        core[v] = d[v];
        maxcore = std::max(maxcore, core[v]);
        if (!live(v)) continue;
        for (uint32_t e = ro[v]; e < ro[v + 1]; ++e) {
            const int32_t u = static_cast<int32_t>(cols[e]);
            if (live(u) && d[u] > d[v]) {  // move u one bucket down
                const int64_t du = d[u];
                const int32_t pu = pos[u], pw = bin[du], w = vert[pw];
                if (u != w) {
                    pos[u] = pw;
                    vert[pu] = w;
                    pos[w] = pu;
                    vert[pw] = u;
                }
                ++bin[du];
                --d[u];
            }
        }
    }
    // output row of each vertex (its rank among the alive ones when only those are written)
    std::vector<int32_t> &row = fs.pos;
    int32_t rows = 0;
    for (int32_t v = 0; v < n; ++v) {
        row[v] = ids ? rows : v;
        rows += live(v) ? 1 : 0;
    }
#pragma omp parallel for schedule(static) if (par)
    for (int32_t v = 0; v < n; ++v) {
        if (!live(v) && ids) continue;
        float *f = out + static_cast<size_t>(row[v]) * 4;
        if (ids) ids[row[v]] = id_of ? id_of[v] : v;
        if (!live(v)) {
            f[0] = f[1] = f[2] = f[3] = 0.0f;
            continue;
        }
        f[0] = static_cast<float>(std::log1p(static_cast<double>(deg[v])));
        f[1] = static_cast<float>(std::log1p(static_cast<double>(two[v])));
        f[2] = static_cast<float>(static_cast<double>(core[v]) / static_cast<double>(std::max<int64_t>(1, maxcore)));
        f[3] = static_cast<float>(deg[v] ? wsum[v] / static_cast<double>(deg[v]) / std::max(wmean, 1e-9) : 0.0);
    }
}
}  // namespace

int mcb_vertex_features(int32_t nodes, const uint32_t *ro, const uint32_t *cols, const int32_t *weights,
                        const uint8_t *alive, float *out) {
    if (nodes <= 0 || !ro || !cols || !weights || !out) return fail(MCB_E_INVALID, "invalid arguments");
    try {
        FeatureScratch fs;
        vertex_features_csr(nodes, ro, cols, weights, alive, out, nullptr, nullptr, fs);
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
    return MCB_OK;
}

// The round construction's state: the 2-core (parallel edges counted) of the graph minus the picks so
// far, kept as a CSR over local ids that is compacted to the surviving vertices whenever fewer than three
// quarters of its vertices are still alive, so that a round costs time in the size of what is left.
struct mcb_fvs_rounds {
    int32_t n = 0;                   // vertices of the input graph
    std::vector<uint32_t> ro, cols;  // local CSR (vertices ascending in input id, row order kept)
    std::vector<int32_t> w;
    std::vector<int32_t> gid;        // local -> input id
    std::vector<int32_t> lid;        // input id -> local id (-1: gone)
    std::vector<uint8_t> alive;      // local
    std::vector<int32_t> d;          // edge ends to alive vertices, per alive local vertex
    std::vector<int32_t> last_ids;   // ids of the last feature rows, in row order
    FeatureScratch scratch;
    int32_t left = 0;

    void kill(int32_t v, std::vector<int32_t> &stack) {
        alive[v] = 0;
        --left;
        stack.push_back(v);
        while (!stack.empty()) {
            const int32_t x = stack.back();
            stack.pop_back();
            for (uint32_t e = ro[x]; e < ro[x + 1]; ++e) {
                const int32_t u = static_cast<int32_t>(cols[e]);
                if (alive[u] && --d[u] <= 1) {
                    alive[u] = 0;
                    --left;
                    stack.push_back(u);
                }
            }
        }
    }

    void compact() {
        const int32_t ln = static_cast<int32_t>(gid.size());
        std::vector<int32_t> map(ln, -1);
        int32_t m = 0;
        for (int32_t v = 0; v < ln; ++v)
            if (alive[v]) map[v] = m++;
        std::vector<uint32_t> nro(static_cast<size_t>(m) + 1, 0), ncols;
        std::vector<int32_t> nw, ngid(m), nd(m);
        for (int32_t v = 0; v < ln; ++v) {
            if (map[v] < 0) {
                lid[gid[v]] = -1;
                continue;
            }
            const int32_t a = map[v];
            for (uint32_t e = ro[v]; e < ro[v + 1]; ++e)
                if (map[cols[e]] >= 0) {
                    ncols.push_back(static_cast<uint32_t>(map[cols[e]]));
                    nw.push_back(w[e]);
                }
            nro[a + 1] = static_cast<uint32_t>(ncols.size());
            ngid[a] = gid[v];
            lid[gid[v]] = a;
            nd[a] = d[v];
        }
        ro.swap(nro);
        cols.swap(ncols);
        w.swap(nw);
        gid.swap(ngid);
        d.swap(nd);
        alive.assign(m, 1);
    }
};

int mcb_fvs_rounds_create(int32_t nodes, const uint32_t *ro, const uint32_t *cols, const int32_t *weights,
                          mcb_fvs_rounds **out) {
    if (nodes <= 0 || !ro || !cols || !weights || !out) return fail(MCB_E_INVALID, "invalid arguments");
    *out = nullptr;
    try {
        std::unique_ptr<mcb_fvs_rounds> h(new mcb_fvs_rounds());
        const int32_t n = nodes;
        h->n = n;
        h->ro.assign(ro, ro + n + 1);
        h->cols.assign(cols, cols + ro[n]);
        h->w.assign(weights, weights + ro[n]);
        for (uint32_t e = 0; e < ro[n]; ++e)
            if (cols[e] >= static_cast<uint32_t>(n)) return fail(MCB_E_INVALID, "column index out of range");
        h->gid.resize(n);
        std::iota(h->gid.begin(), h->gid.end(), 0);
        h->lid = h->gid;
        h->alive.assign(n, 1);
        h->d.resize(n);
        h->left = n;
        std::vector<int32_t> stack;
        for (int32_t v = 0; v < n; ++v) h->d[v] = static_cast<int32_t>(ro[v + 1] - ro[v]);
        for (int32_t v = 0; v < n; ++v)
            if (h->alive[v] && h->d[v] <= 1) h->kill(v, stack);
        if (static_cast<int64_t>(h->left) * 4 < static_cast<int64_t>(n) * 3) h->compact();
        *out = h.release();
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
    return MCB_OK;
}

int32_t mcb_fvs_rounds_left(const mcb_fvs_rounds *h) { return h ? h->left : fail(MCB_E_INVALID, "invalid arguments"); }

int32_t mcb_fvs_rounds_features(mcb_fvs_rounds *h, int32_t cap, int32_t *ids, float *out) {
    if (!h || !ids || !out) return fail(MCB_E_INVALID, "invalid arguments");
    if (cap < h->left) return fail(MCB_E_INVALID, "output capacity below the number of vertices left");
    try {
        vertex_features_csr(static_cast<int32_t>(h->gid.size()), h->ro.data(), h->cols.data(), h->w.data(),
                            h->alive.data(), out, ids, h->gid.data(), h->scratch);
        h->last_ids.assign(ids, ids + h->left);
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
    return h->left;
}

int32_t mcb_fvs_rounds_take_top(mcb_fvs_rounds *h, const float *scores, int32_t count, int32_t k, int32_t *picks) {
    if (!h || !scores || !picks || k < 0) return fail(MCB_E_INVALID, "invalid arguments");
    if (h->last_ids.empty() || count != static_cast<int32_t>(h->last_ids.size()) || count != h->left)
        return fail(MCB_E_STATE, "scores do not belong to the last feature rows");
    k = std::min(k, count);
    try {
        // best first: higher score, NaN last, ties by row (= ascending id)
        std::vector<int32_t> rows(count);
        std::iota(rows.begin(), rows.end(), 0);
        auto before = [&](int32_t a, int32_t b) {
            const float sa = scores[a], sb = scores[b];
            const bool na = std::isnan(sa), nb = std::isnan(sb);
            if (na != nb) return nb;
            if (!na && sa != sb) return sa > sb;
            return a < b;
        };
        std::partial_sort(rows.begin(), rows.begin() + k, rows.end(), before);
        std::vector<int32_t> stack;
        for (int32_t i = 0; i < k; ++i) {
            const int32_t g = h->last_ids[rows[i]];
            picks[i] = g;
            const int32_t v = h->lid[g];
            if (h->alive[v]) h->kill(v, stack);
        }
        h->last_ids.clear();
        if (static_cast<int64_t>(h->left) * 4 < static_cast<int64_t>(h->gid.size()) * 3) h->compact();
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
    return h->left;
}

void mcb_fvs_rounds_free(mcb_fvs_rounds *h) { delete h; }

int mcb_fvs_peel(int32_t nodes, const uint32_t *ro, const uint32_t *cols, uint8_t *alive) {
    if (nodes <= 0 || !ro || !cols || !alive) return fail(MCB_E_INVALID, "invalid arguments");
    try {
        const int32_t n = nodes;
        std::vector<int64_t> d(n, 0);
        std::vector<int32_t> stack;
        for (int32_t v = 0; v < n; ++v) {
            if (!alive[v]) continue;
            for (uint32_t e = ro[v]; e < ro[v + 1]; ++e) d[v] += alive[cols[e]] ? 1 : 0;
            if (d[v] <= 1) stack.push_back(v);
        }
        for (int32_t v : stack) alive[v] = 0;
        while (!stack.empty()) {
            const int32_t v = stack.back();
            stack.pop_back();
            for (uint32_t e = ro[v]; e < ro[v + 1]; ++e) {
                const int32_t u = static_cast<int32_t>(cols[e]);
                if (alive[u] && --d[u] <= 1) {
                    alive[u] = 0;
                    stack.push_back(u);
                }
            }
        }
        int32_t left = 0;
        for (int32_t v = 0; v < n; ++v) left += alive[v] ? 1 : 0;
        return left;
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
}

int mcb_fvs_greedy(int32_t nodes, int32_t directed_edges, const uint32_t *ro, const uint32_t *cols, int32_t *fvs_rank) {
    if (nodes <= 0 || directed_edges < 0 || !ro || !cols || !fvs_rank) return fail(MCB_E_INVALID, "invalid arguments");
    if (ro[nodes] != static_cast<uint32_t>(directed_edges))
        return fail(MCB_E_INVALID, "row offsets do not match the edge count");
    try {
        const int32_t n = nodes;
        std::vector<int64_t> d(n);
        std::vector<uint8_t> alive(n, 1);
        for (int32_t v = 0; v < n; ++v) d[v] = ro[v + 1] - ro[v];
        std::vector<int32_t> stack, picked;
        // removes v and every vertex left with at most one incident edge, transitively
        auto remove = [&](int32_t v0) {
            stack.push_back(v0);
            alive[v0] = 0;
            while (!stack.empty()) {
                const int32_t v = stack.back();
                stack.pop_back();
                for (uint32_t e = ro[v]; e < ro[v + 1]; ++e) {
                    const int32_t u = static_cast<int32_t>(cols[e]);
                    if (!alive[u]) continue;
                    if (--d[u] <= 1) {
                        alive[u] = 0;
                        stack.push_back(u);
                    }
                }
            }
        };
        for (int32_t v = 0; v < n; ++v)
            if (alive[v] && d[v] <= 1) remove(v);
        // max current degree first (ties: lowest id); stale heap entries are skipped
        std::priority_queue<std::pair<int64_t, int32_t>> heap;
        for (int32_t v = 0; v < n; ++v)
            if (alive[v]) heap.push({d[v], -v});
        while (!heap.empty()) {
            const auto [dv, nv] = heap.top();
            heap.pop();
            const int32_t v = -nv;
            if (!alive[v] || dv != d[v]) {
                if (alive[v]) heap.push({d[v], -v});
                continue;
            }
            picked.push_back(v);
            remove(v);
        }
        // minimal subset: reverse-delete trying the vertices outside the pick first, then the picks
        // from the last to the first
        std::vector<uint8_t> in_pick(n, 0);
        for (int32_t v : picked) in_pick[v] = 1;
        std::vector<int32_t> order;
        order.reserve(n);
        for (int32_t v = 0; v < n; ++v)
            if (!in_pick[v]) order.push_back(v);
        for (auto it = picked.rbegin(); it != picked.rend(); ++it) order.push_back(*it);
        return fvs_reverse_delete_impl(n, ro, cols, order.data(), fvs_rank);
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
}

int mcb_problem_build_edges(int32_t nodes, int32_t edges, const int32_t *u, const int32_t *v, const int32_t *w,
                            const mcb_build_options *opt, mcb_problem **out) {
    if (!out || nodes <= 0 || edges < 0 || (edges > 0 && (!u || !v || !w)))
        return fail(MCB_E_INVALID, "invalid arguments");
    *out = nullptr;
    mcb_build_options def;
    mcb_build_options_init(&def);
    if (!opt) opt = &def;
    try {
        InputGraph g;
        g.n = nodes;
        g.rows.reserve(2 * static_cast<size_t>(edges));
        g.cols.reserve(2 * static_cast<size_t>(edges));
        g.w.reserve(2 * static_cast<size_t>(edges));
        int64_t sum = 0;
        for (int32_t i = 0; i < edges; ++i) {
            if (u[i] < 0 || u[i] >= nodes || v[i] < 0 || v[i] >= nodes)
                return fail(MCB_E_INVALID, "edge endpoint out of range");
            g.rows.push_back(static_cast<uint32_t>(u[i]));
            g.cols.push_back(static_cast<uint32_t>(v[i]));
            g.w.push_back(w[i]);
            g.rows.push_back(static_cast<uint32_t>(v[i]));
            g.cols.push_back(static_cast<uint32_t>(u[i]));
            g.w.push_back(w[i]);
            sum += w[i];
        }
        return build_from_input(g, sum, *opt, out);
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
}

// Matrix-Market reader with the acceptance rules of include/FileReader.h:27-67 (coordinate,
// integer or real, symmetric or general; entries "u v w", 1-based).  Values must be integers.
int mcb_problem_build_mtx(const char *path, const mcb_build_options *opt, mcb_problem **out) {
    if (!path || !out) return fail(MCB_E_INVALID, "invalid arguments");
    *out = nullptr;
    FILE *f = std::fopen(path, "r");
    if (!f) return fail(MCB_E_IO, std::string("cannot open ") + path);
    std::string line;
    if (!read_line(f, line)) {
        std::fclose(f);
        return fail(MCB_E_FORMAT, "empty file");
    }
    char b0[64] = {0}, b1[64] = {0}, b2[64] = {0}, b3[64] = {0}, b4[64] = {0};
    if (std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", b0, b1, b2, b3, b4) != 5 ||
        std::string(b0) != "%%MatrixMarket") {
        std::fclose(f);
        return fail(MCB_E_FORMAT, "Could not process Matrix Market banner.");
    }
    const std::string obj = lower(b1), fmt = lower(b2), field = lower(b3), sym = lower(b4);
    if (!(obj == "matrix" && fmt == "coordinate" && (field == "integer" || field == "real") &&
          (sym == "symmetric" || sym == "general"))) {
        std::fclose(f);
        return fail(MCB_E_FORMAT, "Sorry, this application does not support this mtx file.");
    }
    long long M = 0, Ncol = 0, nnz = 0;
    bool got = false;
    while (read_line(f, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '%') continue;
        if (std::sscanf(line.c_str(), "%lld %lld %lld", &M, &Ncol, &nnz) == 3) got = true;
        break;
    }
    if (!got || M <= 0 || nnz < 0 || M > 0x7FFFFFFF || nnz > 0x3FFFFFFF) {
        std::fclose(f);
        return fail(MCB_E_FORMAT, "Couldn't find all 3 parameters");
    }
    std::vector<int32_t> uu(nnz), vv(nnz), ww(nnz);
    for (long long i = 0; i < nnz; ++i) {
        long long a, b;
        double wt;
        if (std::fscanf(f, "%lld %lld %lf", &a, &b, &wt) != 3) {
            std::fclose(f);
            return fail(MCB_E_FORMAT, "truncated edge list");
        }
        if (wt != static_cast<double>(static_cast<int32_t>(wt))) {
            std::fclose(f);
            return fail(MCB_E_FORMAT, "edge weight is not a 32-bit integer");
        }
        uu[i] = static_cast<int32_t>(a - 1);
        vv[i] = static_cast<int32_t>(b - 1);
        ww[i] = static_cast<int32_t>(wt);
    }
    std::fclose(f);
    return mcb_problem_build_edges(static_cast<int32_t>(M), static_cast<int32_t>(nnz), uu.data(), vv.data(),
                                   ww.data(), opt, out);
}

int mcb_problem_get_info(const mcb_problem *p, mcb_problem_info *info) {
    if (!p || !info) return fail(MCB_E_INVALID, "invalid arguments");
    *info = p->info;
    return MCB_OK;
}

int mcb_problem_array(const mcb_problem *p, int which, const void **data, int64_t *count) {
    if (!p || !data || !count) return fail(MCB_E_INVALID, "invalid arguments");
#define MCB_ARR(id, vec)                                      \
    case id:                                                  \
        *data = (p->vec).data();                              \
        *count = static_cast<int64_t>((p->vec).size());       \
        return MCB_OK;
    switch (which) {
        MCB_ARR(MCB_A_ROWS, g.rows)
        MCB_ARR(MCB_A_COLS, g.cols)
        MCB_ARR(MCB_A_WEIGHTS, g.w)
        MCB_ARR(MCB_A_REVERSE, g.rev)
        MCB_ARR(MCB_A_ROW_OFFSETS, g.ro)
        MCB_ARR(MCB_A_FVS_RANK, fvs_rank)
        MCB_ARR(MCB_A_ROOTS, roots)
        MCB_ARR(MCB_A_NT_EDGES, nt_edges)
        MCB_ARR(MCB_A_NT_MAP, nt_map)
        MCB_ARR(MCB_A_CAND_ROOT, cand_root)
        MCB_ARR(MCB_A_CAND_EDGE, cand_edge)
        MCB_ARR(MCB_A_CAND_WEIGHT, cand_weight)
        MCB_ARR(MCB_A_CAND_POS_U, cand_pu)
        MCB_ARR(MCB_A_CAND_POS_V, cand_pv)
        MCB_ARR(MCB_A_CAND_COORD, cand_coord)
        MCB_ARR(MCB_A_TREE_OFFSET, tree_offset)
        MCB_ARR(MCB_A_PARENT_POS, parent_pos)
        MCB_ARR(MCB_A_POS_COORD, pos_coord)
        MCB_ARR(MCB_A_ROOT_CAND_OFFSET, root_cand_offset)
        MCB_ARR(MCB_A_ROOT_CAND_RANK, root_cand_rank)
        MCB_ARR(MCB_A_VERTEX_PARENT, vertex_parent)
        MCB_ARR(MCB_A_POS_VERTEX, pos_vertex)
        default:
            return fail(MCB_E_INVALID, "unknown array id");
    }
#undef MCB_ARR
}

namespace {
struct LevelSorter {  // edge_sorter (include/Dijkstra.h:7-21)
    int edge;
    int level;
};
}  // namespace

int mcb_problem_reference_layout(const mcb_problem *p, int32_t r, uint32_t *rows, uint32_t *cols, int32_t *edge_offset,
                                 uint32_t *nodes_index) {
    if (!p || !rows || !cols || !edge_offset || !nodes_index) return fail(MCB_E_INVALID, "invalid arguments");
    if (p->info.single_cycle || r < 0 || r >= p->info.fvs_size) return fail(MCB_E_INVALID, "root index out of range");
    const Multigraph &g = p->g;
    const int32_t n = g.n;
    const uint32_t src = p->roots[r];
    Workspace ws;
    ws.prepare(n, p->fvs_rank.data());
    dijkstra(g, src, false, ws, ws.std_heap);
    const SpTree &t = ws.t;
    std::vector<LevelSorter> edges;
    edges.reserve(n);
    edges.push_back({-1, 0});
    for (uint32_t v : t.order) edges.push_back({t.v[v].pedge, t.v[v].level});
    std::sort(edges.begin(), edges.end(),
              [](const LevelSorter &a, const LevelSorter &b) { return a.level < b.level; });
    std::memset(rows, 0, sizeof(uint32_t) * (n + 1));
    for (size_t i = 0; i < edges.size(); ++i) {
        rows[edges[i].level]++;
        if (edges[i].edge == -1) {
            nodes_index[src] = static_cast<uint32_t>(i);
            cols[i] = 0xFFFFFFFFu;
            edge_offset[i] = -1;
        } else {
            const uint32_t row = g.rows[edges[i].edge], col = g.cols[edges[i].edge];
            nodes_index[col] = static_cast<uint32_t>(i);
            cols[i] = nodes_index[row];
            edge_offset[i] = edges[i].edge;
        }
    }
    uint32_t prev = 0;
    for (int32_t i = 0; i <= n; ++i) {
        uint32_t cur = rows[i];
        rows[i] = prev;
        prev += cur;
    }
    return MCB_OK;
}

void mcb_problem_free(mcb_problem *p) { delete p; }

int mcb_problem_finish_trees(mcb_problem *p, const mcb_root_trees *t, int num_threads) {
    if (!p || !t || !t->status) return fail(MCB_E_INVALID, "invalid arguments");
    if (!p->trees_pending) return fail(MCB_E_STATE, "the problem was not built with external_trees=1");
    const int32_t F = p->info.fvs_size;
    StageClock clock;
    std::vector<RootOutput> per_root;
    try {
        per_root.resize(F);
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
    const int threads = num_threads > 0 ? num_threads : omp_get_max_threads();
    int err = 0;
#pragma omp parallel num_threads(threads)
    {
        Workspace ws;
#pragma omp for schedule(dynamic, 16)
        for (int32_t r = 0; r < F; ++r) {
            try {
                if (t->status[r] == 0) {
                    const int64_t c0 = t->cand_start[r], k0 = t->kept_start[r];
                    root_from_search(*p, p->roots[r], t->cand_count[r], t->cand_edge + c0, t->cand_weight + c0,
                                     t->kept_count[r], t->kept + 4 * k0, ws, per_root[r]);
                } else {
                    process_root(*p, p->roots[r], true, ws, per_root[r], nullptr);
                }
            } catch (const std::bad_alloc &) {
#pragma omp atomic write
                err = 1;
            }
        }
    }
    if (err) return fail(MCB_E_NOMEM, "out of memory while laying out shortest-path trees");
    clock.mark("trees from search results");
    int rc;
    try {
        rc = assemble(*p, per_root, false, threads, clock);
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory while assembling the trees");
    }
    if (rc) return rc;
    p->trees_pending = false;
    return MCB_OK;
}

int mcb_problem_root_search(const mcb_problem *p, int32_t r, int64_t cand_cap, uint32_t *cand_edge,
                            int32_t *cand_weight, int64_t kept_cap, uint32_t *kept, int32_t *num_cand,
                            int32_t *num_kept) {
    if (!p || !num_cand || !num_kept || r < 0 || r >= p->info.fvs_size || p->roots.empty())
        return fail(MCB_E_INVALID, "invalid arguments");
    Workspace ws;
    std::vector<uint32_t> ce, kv;
    std::vector<int32_t> cw;
    try {
        search_results(*p, p->roots[r], ws, ce, cw, kv);
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
    *num_cand = static_cast<int32_t>(ce.size());
    *num_kept = static_cast<int32_t>(kv.size() / 4);
    if (static_cast<int64_t>(ce.size()) > cand_cap || static_cast<int64_t>(kv.size() / 4) > kept_cap)
        return fail(MCB_E_INVALID, "capacity too small for the root's results");
    if (!ce.empty() && (!cand_edge || !cand_weight)) return fail(MCB_E_INVALID, "invalid arguments");
    if (!kv.empty() && !kept) return fail(MCB_E_INVALID, "invalid arguments");
    std::copy(ce.begin(), ce.end(), cand_edge);
    std::copy(cw.begin(), cw.end(), cand_weight);
    std::copy(kv.begin(), kv.end(), kept);
    return MCB_OK;
}

int mcb_fvs_reverse_delete(int32_t nodes, int32_t directed_edges, const uint32_t *row_offsets, const uint32_t *cols,
                           const int32_t *order, int32_t *fvs_rank) {
    if (nodes <= 0 || directed_edges < 0 || !row_offsets || !cols || !fvs_rank)
        return fail(MCB_E_INVALID, "invalid arguments");
    if (row_offsets[nodes] != static_cast<uint32_t>(directed_edges))
        return fail(MCB_E_INVALID, "row offsets do not match the edge count");
    try {
        return fvs_reverse_delete_impl(nodes, row_offsets, cols, order, fvs_rank);
    } catch (const std::bad_alloc &) {
        return fail(MCB_E_NOMEM, "out of memory");
    }
}

int mcb_induced_has_cycle(int32_t nodes, const uint32_t *row_offsets, const uint32_t *cols, const uint8_t *in_set) {
    if (nodes <= 0 || !row_offsets || !cols || !in_set) return fail(MCB_E_INVALID, "invalid arguments");
    // Count each undirected edge once via its lower-offset direction: an edge joining two vertices
    // already connected closes a cycle.  Parallel edges appear as two offsets per direction.
    UnionFind uf(nodes);
    for (int32_t v = 0; v < nodes; ++v) {
        if (in_set[v]) continue;
        for (uint32_t e = row_offsets[v]; e < row_offsets[v + 1]; ++e) {
            const uint32_t c = cols[e];
            if (in_set[c]) continue;
            if (c == static_cast<uint32_t>(v)) return 1;
            if (c < static_cast<uint32_t>(v)) continue;  // the (c -> v) direction was counted at c
            const int32_t a = uf.find(v), b = uf.find(static_cast<int32_t>(c));
            if (a == b) return 1;
            uf.p[a] = b;
        }
    }
    return 0;
}

int mcb_biconnected_components(int32_t nodes, int32_t edges, const int32_t *u, const int32_t *v,
                               int32_t *edge_component) {
    if (nodes <= 0 || edges < 0 || (edges > 0 && (!u || !v || !edge_component)))
        return fail(MCB_E_INVALID, "invalid arguments");
    // Iterative Hopcroft-Tarjan with an edge stack.
    std::vector<uint32_t> ro(nodes + 1, 0), adj_to(2 * static_cast<size_t>(edges)), adj_id(2 * static_cast<size_t>(edges));
    for (int32_t i = 0; i < edges; ++i) {
        if (u[i] < 0 || u[i] >= nodes || v[i] < 0 || v[i] >= nodes) return fail(MCB_E_INVALID, "endpoint out of range");
        ro[u[i] + 1]++;
        ro[v[i] + 1]++;
    }
    for (int32_t i = 0; i < nodes; ++i) ro[i + 1] += ro[i];
    {
        std::vector<uint32_t> fill(ro.begin(), ro.end() - 1);
        for (int32_t i = 0; i < edges; ++i) {
            adj_to[fill[u[i]]] = static_cast<uint32_t>(v[i]);
            adj_id[fill[u[i]]++] = static_cast<uint32_t>(i);
            adj_to[fill[v[i]]] = static_cast<uint32_t>(u[i]);
            adj_id[fill[v[i]]++] = static_cast<uint32_t>(i);
        }
    }
    std::vector<int64_t> disc(nodes, -1), low(nodes, 0);
    std::vector<uint32_t> estack;
    for (int32_t i = 0; i < edges; ++i) edge_component[i] = -1;
    struct Frame {
        uint32_t vtx, it;
        int64_t pedge;
    };
    std::vector<Frame> st;
    int64_t time = 0;
    int32_t comp = 0;
    for (int32_t s = 0; s < nodes; ++s) {
        if (disc[s] >= 0) continue;
        disc[s] = low[s] = time++;
        st.push_back({static_cast<uint32_t>(s), ro[s], -1});
        while (!st.empty()) {
            Frame &f = st.back();
            const uint32_t x = f.vtx;
            if (f.it < ro[x + 1]) {
                const uint32_t k = f.it++;
                const uint32_t y = adj_to[k], id = adj_id[k];
                if (static_cast<int64_t>(id) == f.pedge) continue;
                if (disc[y] < 0) {
                    estack.push_back(id);
                    disc[y] = low[y] = time++;
                    st.push_back({y, ro[y], static_cast<int64_t>(id)});
                } else if (disc[y] < disc[x]) {
                    estack.push_back(id);
                    low[x] = std::min(low[x], disc[y]);
                }
                continue;
            }
            const int64_t pe = f.pedge;
            st.pop_back();
            if (st.empty()) break;
            const uint32_t p = st.back().vtx;
            low[p] = std::min(low[p], low[x]);
            if (low[x] >= disc[p]) {
                while (!estack.empty()) {
                    const uint32_t id = estack.back();
                    estack.pop_back();
                    edge_component[id] = comp;
                    if (static_cast<int64_t>(id) == pe) break;
                }
                ++comp;
            }
        }
    }
    return comp;
}

}  // extern "C"
