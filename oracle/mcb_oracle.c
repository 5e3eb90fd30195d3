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

/* ---------------------------------------------------------------------------------------------
 * Per-phase checks on the position layout
 * ------------------------------------------------------------------------------------------- */

int oracle_position_parity(int32_t F, const int64_t *tree_offset, const uint32_t *parent_pos,
                           const int32_t *pos_coord, int32_t W, const uint64_t *S, uint8_t *par,
                           int32_t num_threads) {
    if (F <= 0 || !tree_offset || !parent_pos || !pos_coord || W <= 0 || !S || !par) return -3;
    const int threads = num_threads > 0 ? num_threads : omp_get_max_threads();
    int rc = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64)
    for (int32_t r = 0; r < F; ++r) {
        const int64_t a = tree_offset[r], b = tree_offset[r + 1];
        if (b <= a || parent_pos[a] != 0xFFFFFFFFu) {
#pragma omp atomic write
            rc = -3;
            continue;
        }
        par[a] = 0; /* the root's own value (include/work_per_thread.h:207) */
        for (int64_t p = a + 1; p < b; ++p) {
            const uint32_t q = parent_pos[p];
            const int32_t c = pos_coord[p];
            if ((int64_t)q >= p - a || c >= 64 * W) {
#pragma omp atomic write
                rc = -3;
                break;
            }
            /* a child's value is its parent's plus the S bit of the tree edge's coordinate, mod 2
             * (include/work_per_thread.h:189-200); parents precede children in the list */
            par[p] = (uint8_t)(par[a + q] ^ (c >= 0 ? get_bit(S, c) : 0u));
        }
    }
    return rc;
}

int64_t oracle_phase_select(int32_t F, const int64_t *tree_offset, const uint32_t *parent_pos,
                            const int32_t *pos_coord, int64_t K, const uint32_t *cand_root,
                            const uint32_t *cand_pos_u, const uint32_t *cand_pos_v, const int32_t *cand_coord,
                            int32_t W, const uint64_t *S, const uint8_t *taken, int32_t num_threads) {
    if (F <= 0 || !tree_offset || K < 0 || !cand_root || !cand_pos_u || !cand_pos_v || !cand_coord || W <= 0 || !S)
        return -3;
    const int64_t P = tree_offset[F];
    uint8_t *par = (uint8_t *)malloc((size_t)(P > 0 ? P : 1));
    if (!par) return -5;
    int rc = oracle_position_parity(F, tree_offset, parent_pos, pos_coord, W, S, par, num_threads);
    if (rc) {
        free(par);
        return rc;
    }
    int64_t chosen = -1;
    /* the list in rank order; a chosen candidate was erased from it (mcb/mcb_cycle.cpp:279-283) */
    for (int64_t k = 0; k < K; ++k) {
        if (taken && taken[k]) continue;
        const uint32_t root = cand_root[k];
        if (root >= (uint32_t)F) { chosen = -3; break; }
        const int64_t a = tree_offset[root], np = tree_offset[root + 1] - a;
        if ((int64_t)cand_pos_u[k] >= np || (int64_t)cand_pos_v[k] >= np || cand_coord[k] >= 64 * W) {
            chosen = -3;
            break;
        }
        unsigned bit = cand_coord[k] >= 0 ? get_bit(S, cand_coord[k]) : 0u;
        bit ^= par[a + cand_pos_u[k]];
        bit ^= par[a + cand_pos_v[k]];
        if (bit == 1) { chosen = k; break; }
    }
    free(par);
    return chosen;
}

/* The coordinates of rank k's cycle, visited once each: the closing edge, then the edges on the
 * way up from u and from v to the root (the two paths meet only at the root, include/cycles.h:
 * 144-195, so setting and toggling bits agree). */
static int cycle_walk(const int64_t *tree_offset, const uint32_t *parent_pos, const int32_t *pos_coord, int64_t K,
                      const uint32_t *cand_root, const uint32_t *cand_pos_u, const uint32_t *cand_pos_v,
                      const int32_t *cand_coord, int64_t k, void (*visit)(void *, int32_t), void *ctx) {
    if (k < 0 || k >= K) return -3;
    const int64_t a = tree_offset[cand_root[k]], np = tree_offset[cand_root[k] + 1] - a;
    if (cand_coord[k] >= 0) visit(ctx, cand_coord[k]);
    const uint32_t ends[2] = {cand_pos_u[k], cand_pos_v[k]};
    for (int t = 0; t < 2; ++t) {
        uint32_t p = ends[t];
        if ((int64_t)p >= np) return -3;
        int64_t steps = 0;
        while (p != 0) {
            if (pos_coord[a + p] >= 0) visit(ctx, pos_coord[a + p]);
            p = parent_pos[a + p];
            if ((int64_t)p >= np || ++steps > np) return -3;
        }
    }
    return 0;
}

typedef struct {
    uint64_t *C;
    int32_t W;
    int bad;
} cv_ctx;

static void cv_visit(void *ctx, int32_t c) {
    cv_ctx *x = (cv_ctx *)ctx;
    if (c >= 64 * x->W) { x->bad = 1; return; }
    x->C[c >> 6] ^= 1ull << (c & 63);
}

int oracle_cycle_vector(const int64_t *tree_offset, const uint32_t *parent_pos, const int32_t *pos_coord,
                        int64_t K, const uint32_t *cand_root, const uint32_t *cand_pos_u,
                        const uint32_t *cand_pos_v, const int32_t *cand_coord, int64_t k, int32_t W,
                        uint64_t *C) {
    if (!tree_offset || !parent_pos || !pos_coord || !cand_root || !cand_pos_u || !cand_pos_v || !cand_coord ||
        W <= 0 || !C)
        return -3;
    memset(C, 0, sizeof(uint64_t) * (size_t)W);
    cv_ctx x = {C, W, 0};
    int rc = cycle_walk(tree_offset, parent_pos, pos_coord, K, cand_root, cand_pos_u, cand_pos_v, cand_coord, k,
                        cv_visit, &x);
    return rc ? rc : (x.bad ? -3 : 0);
}

typedef struct {
    uint64_t *Y;
    uint64_t r;
    int32_t N;
    int bad;
} orth_ctx;

static void orth_visit(void *ctx, int32_t c) {
    orth_ctx *x = (orth_ctx *)ctx;
    if (c >= x->N) { x->bad = 1; return; }
    x->Y[c] ^= x->r;
}

int oracle_orthogonality(const int64_t *tree_offset, const uint32_t *parent_pos, const int32_t *pos_coord,
                         int64_t K, const uint32_t *cand_root, const uint32_t *cand_pos_u,
                         const uint32_t *cand_pos_v, const int32_t *cand_coord, int32_t N,
                         const int32_t *selected, const uint64_t *r, int32_t nsamp, const int32_t *samp,
                         int32_t W, const uint64_t *S_samp, uint64_t *out) {
    if (!tree_offset || !parent_pos || !pos_coord || !cand_root || !cand_pos_u || !cand_pos_v || !cand_coord ||
        N <= 0 || !selected || !r || nsamp < 0 || (nsamp > 0 && (!samp || !S_samp || !out)) || W * 64 < N)
        return -3;
    /* Y[c] = XOR of r[i] over the phases i done so far whose cycle vector holds coordinate c; then
     * bit t of XOR_{c in S_j} Y[c] is XOR over those i with bit t of r[i] of <C_i, S_j>. */
    uint64_t *Y = (uint64_t *)calloc((size_t)N, sizeof(uint64_t));
    if (!Y) return -5;
    int rc = 0;
    int32_t i = 0;
    for (int32_t s = 0; s < nsamp && !rc; ++s) {
        const int32_t j = samp[s];
        if (j < 0 || j >= N || (s > 0 && j < samp[s - 1])) { rc = -3; break; }
        for (; i < j && !rc; ++i) {
            orth_ctx x = {Y, r[i], N, 0};
            rc = cycle_walk(tree_offset, parent_pos, pos_coord, K, cand_root, cand_pos_u, cand_pos_v, cand_coord,
                            selected[i], orth_visit, &x);
            if (!rc && x.bad) rc = -3;
        }
        if (rc) break;
        const uint64_t *Sj = S_samp + (size_t)s * W;
        uint64_t acc = 0;
        for (int32_t w = 0; w < W; ++w) {
            uint64_t bits = Sj[w];
            while (bits) {
                const int32_t c = w * 64 + __builtin_ctzll(bits);
                bits &= bits - 1;
                if (c >= N) { rc = -3; break; }
                acc ^= Y[c];
            }
        }
        out[s] = acc;
    }
    free(Y);
    return rc;
}

/* ---------------------------------------------------------------------------------------------
 * One timed phase (the CPU baseline)
 * ------------------------------------------------------------------------------------------- */

int64_t oracle_phase_timed(int32_t n, int32_t F, const uint32_t *roots, const uint32_t *rows, const uint32_t *cols,
                           const int32_t *vertex_parent, const int32_t *nt_map, int32_t N, int64_t K,
                           const uint32_t *cand_root, const uint32_t *cand_edge, const int32_t *root_sample,
                           int32_t nsample, int32_t W, const uint64_t *S_e, const uint8_t *taken, uint8_t *P,
                           uint64_t *later, int64_t nlater, int32_t num_threads, double *t) {
    if (n <= 0 || F <= 0 || N <= 0 || K <= 0 || !roots || !rows || !cols || !vertex_parent || !nt_map || !cand_root ||
        !cand_edge || nsample < 0 || (nsample > 0 && !root_sample) || W != (N + 63) / 64 || !S_e || !P || nlater < 0 ||
        (nlater > 0 && !later) || !t)
        return -3;
    const int threads = num_threads > 0 ? num_threads : omp_get_max_threads();
    for (int32_t i = 0; i < nsample; ++i)
        if (root_sample[i] < 0 || root_sample[i] >= F) return -3;
    /* the sampled roots' children lists, built untimed (the reference keeps its trees across phases) */
    children_t *ch = (children_t *)calloc((size_t)(nsample > 0 ? nsample : 1), sizeof(children_t));
    if (!ch) return -5;
    int rc = 0;
    for (int32_t i = 0; i < nsample && !rc; ++i)
        if (build_children(n, rows, cols, vertex_parent + (size_t)root_sample[i] * n, &ch[i])) rc = -5;
    int64_t chosen = -1;
    if (rc) goto done;

    double t0 = omp_get_wtime();
#pragma omp parallel num_threads(threads)
    {
        int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
#pragma omp for schedule(dynamic, 8)
        for (int32_t i = 0; i < nsample; ++i) {
            const int32_t r = root_sample[i];
            if (queue) parity_one_root(n, roots[r], &ch[i], nt_map, S_e, P + (size_t)r * n, queue);
        }
        if (!queue) {
#pragma omp atomic write
            rc = -5;
        }
        free(queue);
    }
    t[0] = omp_get_wtime() - t0;
    if (rc) goto done;

    t0 = omp_get_wtime();
    for (int64_t k = 0; k < K; ++k) {
        if (taken && taken[k]) continue;
        const uint32_t edge = cand_edge[k];
        const uint8_t *Pr = P + (size_t)cand_root[k] * n;
        unsigned bit = 0;
        if (nt_map[edge] >= 0) bit = get_bit(S_e, nt_map[edge]);
        bit = (bit + Pr[rows[edge]]) % 2;
        bit = (bit + Pr[cols[edge]]) % 2;
        if (bit == 1) { chosen = k; break; }
    }
    t[1] = omp_get_wtime() - t0;
    if (chosen < 0) { rc = -8; goto done; }

    t0 = omp_get_wtime();
    {
        uint64_t *C = (uint64_t *)calloc((size_t)W, sizeof(uint64_t));
        if (!C) { rc = -5; goto done; }
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
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t j = 0; j < nlater; ++j) {
            uint64_t *Sj = later + (size_t)j * W;
            if (dot_product(C, Sj, W) == 1)
                for (int32_t i = 0; i < W; ++i) Sj[i] ^= S_e[i];
        }
        free(C);
    }
    t[2] = omp_get_wtime() - t0;
done:
    for (int32_t i = 0; i < nsample; ++i) free_children(&ch[i]);
    free(ch);
    return rc ? rc : chosen;
}
