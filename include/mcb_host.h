/*
 * mcb_host.h - C ABI of the host-side preprocessing library (libmcb_host.so).
 *
 * The host library turns an undirected integer-weighted graph into the inputs of the
 * de Pina support-vector phase loop: the reduced multigraph, the feedback-vertex-set
 * roots, the spanning-tree coordinates of the cycle space, one shortest-path tree per
 * root laid out in BFS-level order, and the candidate cycles in their global rank order.
 * It replaces the host half of the reference's driver
 *   mcb/mcb_cycle_deg_two_remove.cpp:68-267   (GPU driver: read, reduce, FVS, trees, candidates)
 *   mcb/mcb_cycle.cpp:76-203                  (CPU baseline: same without degree-two removal)
 * No function here touches a GPU; every call is thread-safe on distinct problems.
 *
 * All functions return MCB_OK (0) or a negative MCB_E* code; mcb_host_last_error()
 * gives a message for the calling thread.
 */
#ifndef MCB_HOST_H
#define MCB_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MCB_OK = 0,
    MCB_E_IO = -1,          /* file missing or unreadable (FileReader.h:29-32 exits 1) */
    MCB_E_FORMAT = -2,      /* unsupported Matrix-Market banner/size line (FileReader.h:34-50) */
    MCB_E_INVALID = -3,     /* bad argument */
    MCB_E_GRAPH = -4,       /* graph the path cannot handle (disconnected, no vertex of degree > 2 ...) */
    MCB_E_NOMEM = -5,
    MCB_E_DEVICE = -6,      /* HIP runtime error (device library only) */
    MCB_E_STATE = -7,       /* call out of order */
    MCB_E_NOT_FOUND = -8    /* no odd candidate in a phase (cannot happen for a valid candidate set) */
};

typedef struct mcb_problem mcb_problem;

/* Options of mcb_problem_build. */
typedef struct {
    int reduce_degree_two;  /* 1: contract degree-two chains and root the spanning tree at the first
                               vertex of degree > 2 (mcb_cycle_deg_two_remove.cpp, mcb_non_cpu_rd.cpp);
                               0: keep the graph and root it at vertex 0 (mcb_cycle.cpp:110-113) */
    int prune_trees;        /* 1: keep in each root's position layout only the root-to-endpoint paths
                               of that root's candidates (exact: parity is needed only at endpoints) */
    int keep_vertex_trees;  /* 1: also keep, per root, the parent edge of every vertex (F*n' ints),
                               the input of the CPU phase loop and of the reference tree layout */
    int num_threads;        /* OpenMP threads for the per-root work (<=0: library default) */
    const int32_t *fvs_order; /* optional: vertex ids of the reduced graph in the order the reverse-delete
                               pass tries them (NULL: the reference order, degree asc then id desc) */
    int fvs_order_len;
    int stop_after_fvs;     /* 1: build only the reduced multigraph and its FVS (the arrays up to
                               MCB_A_ROOTS; info through fvs_size and num_non_tree) */
    int external_trees;     /* 1: build up to the coordinates (the arrays up to MCB_A_NT_MAP) and leave the
                               shortest-path trees to mcb_problem_finish_trees, fed by a search elsewhere
                               (the device search of mcb_hip.h); needs prune_trees=1, keep_vertex_trees=0 */
} mcb_build_options;

void mcb_build_options_init(mcb_build_options *opt);

/* Scalar facts about a built problem. */
typedef struct {
    int32_t input_nodes;        /* n of the input file */
    int32_t input_edges;        /* m of the input file (undirected) */
    int32_t single_cycle;       /* 1 if every vertex has degree two (mcb_cycle_deg_two_remove.cpp:104-113) */
    int32_t nodes_removed;      /* degree-two vertices contracted away */
    int32_t source_vertex;      /* spanning-tree root (an id of the reduced graph, see O3) */
    int32_t nodes;              /* n' of the reduced multigraph */
    int32_t directed_edges;     /* 2m' (both directions of every edge) */
    int32_t num_non_tree;       /* N = m' - n' + 1, the cycle-space dimension */
    int32_t words;              /* W = ceil(N / 64) */
    int32_t fvs_size;           /* F, number of roots */
    int64_t num_candidates;     /* K */
    int64_t positions;          /* total positions over all root layouts */
    int32_t max_positions;      /* largest single root layout */
    int32_t max_depth;          /* deepest level over all root layouts */
    int64_t input_total_weight; /* sum of input edge weights (int64) */
} mcb_problem_info;

/* Arrays a problem exposes (mcb_problem_array). Element type in brackets, length after it. */
enum {
    MCB_A_ROWS = 0,         /* [u32, 2m']  reduced multigraph, CSR order by (row, col) */
    MCB_A_COLS,             /* [u32, 2m'] */
    MCB_A_WEIGHTS,          /* [i32, 2m'] */
    MCB_A_REVERSE,          /* [u32, 2m']  offset of the opposite direction */
    MCB_A_ROW_OFFSETS,      /* [u32, n'+1] */
    MCB_A_FVS_RANK,         /* [i32, n']   root index of a vertex, -1 if not a root (FVS.cpp:216-227) */
    MCB_A_ROOTS,            /* [u32, F]    vertex of each root index */
    MCB_A_NT_EDGES,         /* [u32, N]    edge offset (row<col direction) of each coordinate */
    MCB_A_NT_MAP,           /* [i32, 2m']  coordinate of an edge offset, both directions, -1 for tree edges */
    MCB_A_CAND_ROOT,        /* [u32, K]    root index of each candidate, in rank order */
    MCB_A_CAND_EDGE,        /* [u32, K]    non-tree edge offset of each candidate */
    MCB_A_CAND_WEIGHT,      /* [i32, K]    cycle weight d(u)+d(v)+w(e) */
    MCB_A_CAND_POS_U,       /* [u32, K]    position of rows[edge] in the root's layout */
    MCB_A_CAND_POS_V,       /* [u32, K]    position of cols[edge] in the root's layout */
    MCB_A_CAND_COORD,       /* [i32, K]    coordinate of the candidate's own edge or -1 */
    MCB_A_TREE_OFFSET,      /* [i64, F+1]  first position of each root's layout */
    MCB_A_PARENT_POS,       /* [u32, P]    parent position (root: 0xFFFFFFFF), relative to the root's layout */
    MCB_A_POS_COORD,        /* [i32, P]    coordinate of the edge to the parent, -1 if a tree edge or root */
    MCB_A_ROOT_CAND_OFFSET, /* [i64, F+1]  candidates of each root, ranks ascending */
    MCB_A_ROOT_CAND_RANK,   /* [u32, K] */
    MCB_A_VERTEX_PARENT,    /* [i32, F*n'] parent edge of each vertex per root (-1 at the root); keep_vertex_trees */
    MCB_A_POS_VERTEX,       /* [u32, P]    vertex at each position (kept with keep_vertex_trees) */
    MCB_A_COUNT
};

int mcb_problem_build_mtx(const char *path, const mcb_build_options *opt, mcb_problem **out);

/* Edges as 0-based endpoint pairs; each undirected edge once. */
int mcb_problem_build_edges(int32_t nodes, int32_t edges, const int32_t *u, const int32_t *v,
                            const int32_t *w, const mcb_build_options *opt, mcb_problem **out);

int mcb_problem_get_info(const mcb_problem *p, mcb_problem_info *info);

/* Pointer to an internal array (valid until mcb_problem_free) and its element count. */
int mcb_problem_array(const mcb_problem *p, int which, const void **data, int64_t *count);

/* The reference's per-root tree layout for root index r, exactly as
 * dijkstra::fill_tree_edges (include/Dijkstra.h:151-203) writes it: rows (level offsets, n'+1),
 * cols (parent position, root 0xFFFFFFFF), edge_offset (edge id, root -1), nodes_index (vertex -> position).
 * Recomputes that root's shortest-path tree. */
int mcb_problem_reference_layout(const mcb_problem *p, int32_t r, uint32_t *rows, uint32_t *cols,
                                 int32_t *edge_offset, uint32_t *nodes_index);

void mcb_problem_free(mcb_problem *p);

/* Per-root results of the stopped shortest-path searches (include/Dijkstra.h:78-127 with the candidate
 * filter of include/Dijkstra.h:205-254), in the form a device search hands them over.  For root index r
 * with status[r] == 0:
 *   cand_edge/cand_weight[cand_start[r] .. + cand_count[r]]  the accepted candidate edges (offset of the
 *       direction listed first, include/Dijkstra.h:135-148) and their cycle weights d(u)+d(v)+w(e), in any
 *       order;
 *   kept[4*(kept_start[r] .. + kept_count[r])]  the vertices of the pruned layout other than the root (the
 *       candidate endpoints and all their tree ancestors), each as four 32-bit words (vertex, parent edge
 *       offset, hop level, finalisation index with the root at 0), in any order.
 * A root with status[r] != 0 is searched by the host library itself. */
typedef struct {
    const int32_t *status;
    const int64_t *cand_start;
    const int32_t *cand_count;
    const uint32_t *cand_edge;
    const int32_t *cand_weight;
    const int64_t *kept_start;
    const int32_t *kept_count;
    const uint32_t *kept;
} mcb_root_trees;

/* Completes a problem built with external_trees=1 from per-root search results: candidate lists in the
 * reference's order and rank (mcb/mcb_cycle_deg_two_remove.cpp:244-258), pruned BFS-level layouts.  The
 * result is identical to building the problem with the trees searched on the host. */
int mcb_problem_finish_trees(mcb_problem *p, const mcb_root_trees *trees, int num_threads);

/* The host's own stopped search of root index r in the mcb_root_trees form (a problem built with
 * external_trees=1 or prune_trees=1).  Writes at most cand_cap candidates and kept_cap kept vertices
 * (4 words each) and their counts; returns MCB_E_INVALID, with the counts set, when a capacity is short. */
int mcb_problem_root_search(const mcb_problem *p, int32_t r, int64_t cand_cap, uint32_t *cand_edge,
                            int32_t *cand_weight, int64_t kept_cap, uint32_t *kept, int32_t *num_cand,
                            int32_t *num_kept);

/* Reverse-delete feedback vertex set of the reduced multigraph (FVS.cpp:166-227), given the order in
 * which vertices are tried (NULL: degree asc, id desc).  Writes rank per vertex or -1; returns F or <0. */
int mcb_fvs_reverse_delete(int32_t nodes, int32_t directed_edges, const uint32_t *row_offsets,
                           const uint32_t *cols, const int32_t *order, int32_t *fvs_rank);

/* 1 if the multigraph induced on vertices with in_set[v]==0 has a cycle (parallel edges count),
 * as FVS::test_fvs (FVS.cpp:145-164); 0 if it is a forest. */
int mcb_induced_has_cycle(int32_t nodes, const uint32_t *row_offsets, const uint32_t *cols,
                          const uint8_t *in_set);

/* Learned-FVS inputs (no counterpart in the reference; they feed the vertex scorer whose picks are
 * passed as mcb_build_options.fvs_order).  Per vertex of the multigraph induced by the vertices with
 * alive[v] != 0 (alive NULL: all), 4 floats: log(1 + degree), log(1 + sum of the neighbours' degrees),
 * core number / largest core number, and mean incident weight / mean weight over all edge ends; zeros
 * for the other vertices.  Parallel edges count with multiplicity. */
int mcb_vertex_features(int32_t nodes, const uint32_t *row_offsets, const uint32_t *cols, const int32_t *weights,
                        const uint8_t *alive, float *out);

/* Clears alive[v] for every vertex left with at most one edge to alive vertices, transitively (such
 * vertices lie on no cycle of the alive subgraph).  Returns the number still alive or <0. */
int mcb_fvs_peel(int32_t nodes, const uint32_t *row_offsets, const uint32_t *cols, uint8_t *alive);

/* The learned construction's rounds in one native state (no counterpart in the reference): the vertices
 * left are those of the 2-core (mcb_fvs_peel) of the graph minus the picks so far.  Each round reads the
 * feature rows of the vertices left (mcb_vertex_features of that subgraph, rows in ascending vertex id,
 * with the ids), scores them, and takes the k best (higher score first, NaN last, ties to the lower id);
 * the picks are removed and the rest peeled again.  A round costs time in the size of what is left
 * (the state compacts itself), not in the size of the graph. */
typedef struct mcb_fvs_rounds mcb_fvs_rounds;
int mcb_fvs_rounds_create(int32_t nodes, const uint32_t *row_offsets, const uint32_t *cols, const int32_t *weights,
                          mcb_fvs_rounds **out);
/* Number of vertices left. */
int32_t mcb_fvs_rounds_left(const mcb_fvs_rounds *h);
/* Writes the ids (ascending) and 4-float feature rows of the vertices left; cap is the row capacity.
 * Returns the number of rows or <0. */
int32_t mcb_fvs_rounds_features(mcb_fvs_rounds *h, int32_t cap, int32_t *ids, float *rows);
/* Takes the k best of the last feature rows by score (count = that call's row count), writes their ids
 * best first to picks[0..min(k, count)), removes them and peels.  Returns the number left or <0. */
int32_t mcb_fvs_rounds_take_top(mcb_fvs_rounds *h, const float *scores, int32_t count, int32_t k, int32_t *picks);
void mcb_fvs_rounds_free(mcb_fvs_rounds *h);

/* A smaller FVS than the reference order's, used as the training target of the vertex scorer: vertices
 * left with at most one edge are peeled away, then the vertex of largest remaining degree (lowest id
 * on ties) is taken and peeled, until nothing is left; the picks are then made minimal by the
 * reverse-delete pass (the other vertices first, then the picks last to first).  Writes rank per vertex
 * or -1 like mcb_fvs_reverse_delete; returns F or <0. */
int mcb_fvs_greedy(int32_t nodes, int32_t directed_edges, const uint32_t *row_offsets, const uint32_t *cols,
                   int32_t *fvs_rank);

/* Biconnected components of a simple undirected graph (edge list, 0-based).  Writes, per edge, its
 * component id (components numbered in order of completion); returns the count or <0. */
int mcb_biconnected_components(int32_t nodes, int32_t edges, const int32_t *u, const int32_t *v,
                               int32_t *edge_component);

const char *mcb_host_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MCB_HOST_H */
