/*
 * mcb_oracle.c - CPU restatement of the reference's de Pina phase loop (test infrastructure only;
 * see mcb_oracle.h).  Plain C11 + OpenMP, no dependency on the product libraries.
 *
 * Pinned by: the reference's own trace drivers (per-phase S_e, chosen cycle, C_e on tiny graphs;
 * chosen-cycle sequences on medium graphs), its CSV totals at full size, brute-force minimum cycle
 * bases on tiny graphs and networkx's minimum_cycle_basis weights (tests/test_oracle_golden.py).
 */
#include "mcb_oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* bit_vector::get_bit (include/bit_vector.h:152-158). */
static inline unsigned get_bit(const uint64_t *s, int32_t pos) {
    return (unsigned)((s[pos >> 6] >> (pos & 63)) & 1ull);
}

/* bit_vector::set_bit with val = true (include/bit_vector.h:118-125). */
static inline void set_bit(uint64_t *s, int32_t pos) { s[pos >> 6] |= 1ull << (pos & 63); }

/* Children of every vertex in one root's tree, as the CSR that csr_multi_graph::fill_tree_edges
 * (include/CsrGraphMulti.h:180-215) builds from the tree edges.  child_edge holds the edge offset
 * (parent -> child) of each child. */
typedef struct {
    int32_t *off;        /* n+1 */
    int32_t *child;      /* n-1 */
    int32_t *child_edge; /* n-1 */
} children_t;

static int build_children(int32_t n, const uint32_t *rows, const uint32_t *cols, const int32_t *vp,
                          children_t *c) {
    c->off = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    c->child = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 1 ? n - 1 : 1));
    c->child_edge = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 1 ? n - 1 : 1));
    if (!c->off || !c->child || !c->child_edge) return -5;
    for (int32_t v = 0; v < n; ++v)
        if (vp[v] >= 0) c->off[rows[vp[v]] + 1]++;
    for (int32_t v = 0; v < n; ++v) c->off[v + 1] += c->off[v];
    int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!fill) return -5;
    memcpy(fill, c->off, sizeof(int32_t) * (size_t)n);
    for (int32_t v = 0; v < n; ++v) {
        if (vp[v] < 0) continue;
        int32_t p = (int32_t)rows[vp[v]];
        if ((int32_t)cols[vp[v]] != v) { free(fill); return -3; }
        c->child[fill[p]] = v;
        c->child_edge[fill[p]++] = vp[v];
    }
    free(fill);
    return 0;
}

static void free_children(children_t *c) {
    free(c->off);
    free(c->child);
    free(c->child_edge);
}

/* worker_thread::precompute_supportVec for one root (include/work_per_thread.h:154-210): BFS from
 * the root; a child's value is its parent's plus the S bit of the tree edge's coordinate, mod 2. */
static void parity_one_root(int32_t n, uint32_t src, const children_t *c, const int32_t *nt_map,
                            const uint64_t *S, uint8_t *P, int32_t *queue) {
    int32_t head = 0, tail = 0;
    P[src] = 0;
    queue[tail++] = (int32_t)src;
    while (head < tail) {
        int32_t top = queue[head++];
        for (int32_t j = c->off[top]; j < c->off[top + 1]; ++j) {
            int32_t column = c->child[j];
            queue[tail++] = column;
            int32_t edge = c->child_edge[j];
            if (nt_map[edge] >= 0)
                P[column] = (uint8_t)((P[top] + get_bit(S, nt_map[edge])) % 2);
            else
                P[column] = P[top];
        }
    }
    P[src] = 0;
    (void)n;
}

int oracle_parity(int32_t n, int32_t F, const uint32_t *roots, const uint32_t *rows, const uint32_t *cols,
                  const int32_t *vertex_parent, const int32_t *nt_map, int32_t W, const uint64_t *S,
                  uint8_t *P, int32_t num_threads) {
    if (n <= 0 || F <= 0 || !roots || !rows || !cols || !vertex_parent || !nt_map || W <= 0 || !S || !P) return -3;
    int rc = 0;
#pragma omp parallel num_threads(num_threads > 0 ? num_threads : omp_get_max_threads())
    {
        int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
#pragma omp for schedule(dynamic, 8)
        for (int32_t r = 0; r < F; ++r) {
            children_t c = {0};
            if (!queue || build_children(n, rows, cols, vertex_parent + (size_t)r * n, &c)) {
#pragma omp atomic write
                rc = -5;
            } else {
                parity_one_root(n, roots[r], &c, nt_map, S, P + (size_t)r * n, queue);
            }
            free_children(&c);
        }
        free(queue);
    }
    return rc;
}

/* bit_vector::get_and_numbers (include/bit_vector.h:73-84): parity of the AND of two words (the
 * reference clears the lowest set bit until none is left; the parity is the same). */
static inline unsigned and_parity(uint64_t a, uint64_t b) { return (unsigned)__builtin_parityll(a & b); }

/* bit_vector::dot_product (include/bit_vector.h:134-141). */
static unsigned dot_product(const uint64_t *a, const uint64_t *b, int32_t W) {
    unsigned val = 0;
    for (int32_t i = 0; i < W; ++i) val ^= and_parity(a[i], b[i]);
    return val;
}

int oracle_run(int32_t n, int32_t F, const uint32_t *roots, const uint32_t *rows, const uint32_t *cols,
               const int32_t *vertex_parent, const int32_t *nt_map, int32_t N, int64_t K,
               const uint32_t *cand_root, const uint32_t *cand_edge, int32_t phases, int32_t num_threads,
               int32_t *selected, uint64_t *S_out, uint64_t *C_out) {
    if (n <= 0 || F <= 0 || N <= 0 || K <= 0 || phases < 0 || phases > N || !roots || !rows || !cols ||
        !vertex_parent || !nt_map || !cand_root || !cand_edge || !selected)
        return -3;
    const int32_t W = (N + 63) / 64;
    const int threads = num_threads > 0 ? num_threads : omp_get_max_threads();
    int rc = 0;

    /* support_vectors[i] = e_i (mcb/mcb_cycle.cpp:208-212) */
    uint64_t *S = (uint64_t *)calloc((size_t)N * W, sizeof(uint64_t));
    uint64_t *C = (uint64_t *)calloc((size_t)W, sizeof(uint64_t));
    uint8_t *P = (uint8_t *)malloc((size_t)F * n);
    uint8_t *taken = (uint8_t *)calloc((size_t)K, 1);
    children_t *ch = (children_t *)calloc((size_t)F, sizeof(children_t));
    if (!S || !C || !P || !taken || !ch) { rc = -5; goto done; }
    for (int32_t i = 0; i < N; ++i) set_bit(S + (size_t)i * W, i);
    for (int32_t r = 0; r < F; ++r)
        if (build_children(n, rows, cols, vertex_parent + (size_t)r * n, &ch[r])) { rc = -5; goto done; }

    for (int32_t e = 0; e < phases; ++e) {
        const uint64_t *Se = S + (size_t)e * W;
        /* precompute (mcb/mcb_cycle.cpp:237-241) */
#pragma omp parallel num_threads(threads)
        {
            int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
#pragma omp for schedule(dynamic, 8)
            for (int32_t r = 0; r < F; ++r) parity_one_root(n, roots[r], &ch[r], nt_map, Se, P + (size_t)r * n, queue);
            free(queue);
        }
        /* inspection: first candidate in list order with odd parity (mcb/mcb_cycle.cpp:254-285);
         * a chosen candidate is erased from the list, here marked taken. */
        int64_t chosen = -1;
        for (int64_t k = 0; k < K; ++k) {
            if (taken[k]) continue;
            const uint32_t edge = cand_edge[k];
            const uint8_t *Pr = P + (size_t)cand_root[k] * n;
            unsigned bit = 0;
            if (nt_map[edge] >= 0) bit = get_bit(Se, nt_map[edge]);
            bit = (bit + Pr[rows[edge]]) % 2;
            bit = (bit + Pr[cols[edge]]) % 2;
            if (bit == 1) { chosen = k; break; }
        }
        if (chosen < 0) { rc = -8; goto done; }
        taken[chosen] = 1;
        selected[e] = (int32_t)chosen;

        /* cycle::get_cycle_vector (include/cycles.h:83-134) */
        memset(C, 0, sizeof(uint64_t) * (size_t)W);
        {
            const uint32_t edge = cand_edge[chosen];
            const int32_t *vp = vertex_parent + (size_t)cand_root[chosen] * n;
            uint32_t row = rows[edge], col = cols[edge];
            if (nt_map[edge] >= 0) set_bit(C, nt_map[edge]);
            while (vp[row] != -1) {
                uint32_t eo = (uint32_t)vp[row];
                if (nt_map[eo] >= 0) set_bit(C, nt_map[eo]);
                row = (rows[eo] != row) ? rows[eo] : cols[eo];
            }
            while (vp[col] != -1) {
                uint32_t eo = (uint32_t)vp[col];
                if (nt_map[eo] >= 0) set_bit(C, nt_map[eo]);
                col = (rows[eo] != col) ? rows[eo] : cols[eo];
            }
        }
        if (C_out) memcpy(C_out + (size_t)e * W, C, sizeof(uint64_t) * (size_t)W);

        /* independence test (mcb/mcb_cycle.cpp:296-301) */
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int32_t j = e + 1; j < N; ++j) {
            uint64_t *Sj = S + (size_t)j * W;
            if (dot_product(C, Sj, W) == 1)
                for (int32_t i = 0; i < W; ++i) Sj[i] ^= Se[i];
        }
    }
    if (S_out) memcpy(S_out, S, sizeof(uint64_t) * (size_t)N * W);
done:
    if (ch) {
        for (int32_t r = 0; r < F; ++r) free_children(&ch[r]);
        free(ch);
    }
    free(S);
    free(C);
    free(P);
    free(taken);
    return rc;
}
