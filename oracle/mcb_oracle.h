/*
 * mcb_oracle.h - CPU restatement of the reference's de Pina phase loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared against; only the
 * test suite, the smoke check and the benchmark's CPU-baseline leg may load it.  The product
 * never calls it.
 *
 * It follows the CPU baseline driver mcb/mcb_cycle.cpp:226-305 line for line on the
 * vertex-indexed tree representation:
 *   parity      worker_thread::precompute_supportVec   include/work_per_thread.h:154-210
 *   search      first odd candidate in list order      mcb/mcb_cycle.cpp:254-285
 *   cycle       cycle::get_cycle_vector                include/cycles.h:83-134
 *   update      dot product / XOR over j > e           mcb/mcb_cycle.cpp:296-301,
 *                                                      include/bit_vector.h:73-84,127-141
 * Inputs are the preprocessed problem (reduced multigraph endpoints, coordinate map, per-root
 * parent edges, candidates in rank order).  Bit vectors are uint64 words, bit i in word i/64
 * at position i%64 (include/bit_vector.h:118-125).
 */
#ifndef MCB_ORACLE_H
#define MCB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-vertex parity of every root's tree for one support vector S (W words):
 * P[r*n + v] = XOR of S bits of the coordinates on the tree path root(r) -> v. */
int oracle_parity(int32_t n, int32_t F, const uint32_t *roots, const uint32_t *rows, const uint32_t *cols,
                  const int32_t *vertex_parent, const int32_t *nt_map, int32_t W, const uint64_t *S,
                  uint8_t *P, int32_t num_threads);

/* Runs phases 0..phases-1 (phases <= N).  selected[e] = rank of the candidate chosen in phase e.
 * S_out (N*W words, optional) receives every support vector as used in its phase (row e = S_e).
 * C_out (phases*W words, optional) receives the cycle vector of every phase.
 * Returns 0, or -8 when a phase finds no odd candidate, or -3 on bad arguments. */
int oracle_run(int32_t n, int32_t F, const uint32_t *roots, const uint32_t *rows, const uint32_t *cols,
               const int32_t *vertex_parent, const int32_t *nt_map, int32_t N, int64_t K,
               const uint32_t *cand_root, const uint32_t *cand_edge, int32_t phases, int32_t num_threads,
               int32_t *selected, uint64_t *S_out, uint64_t *C_out);

/* One phase e of the loop of mcb/mcb_cycle.cpp:237-301 started from a given S_e, timed in parts for
 * the CPU baseline (seconds in t[0..2]):
 *   t[0] precompute: per-vertex parity BFS (include/work_per_thread.h:154-210) of the roots listed in
 *        root_sample (nsample of them), on num_threads threads;
 *   t[1] inspection: the list scan from rank 0 skipping taken ranks (mcb/mcb_cycle.cpp:254-285) over
 *        P (F x n bytes, per-vertex parity of every root, filled by the caller for this S_e);
 *   t[2] independence: the cycle vector of the chosen rank (include/cycles.h:83-134) and, for each of
 *        the nlater vectors in later (nlater x W words), the product with it over all W words and the
 *        XOR of S_e where the product is odd (mcb/mcb_cycle.cpp:296-301).
 * The parity of the sampled roots is written into P as well.  Returns the chosen rank, -8 when no
 * candidate is odd, -3 on bad arguments, -5 when out of memory. */
int64_t oracle_phase_timed(int32_t n, int32_t F, const uint32_t *roots, const uint32_t *rows, const uint32_t *cols,
                           const int32_t *vertex_parent, const int32_t *nt_map, int32_t N, int64_t K,
                           const uint32_t *cand_root, const uint32_t *cand_edge, const int32_t *root_sample,
                           int32_t nsample, int32_t W, const uint64_t *S_e, const uint8_t *taken, uint8_t *P,
                           uint64_t *later, int64_t nlater, int32_t num_threads, double *t);

/* ---------------------------------------------------------------------------------------------
 * Per-phase checks on the position layout (no F x n vertex rows, so they reach the largest graphs).
 *
 * A root r's tree is listed as positions tree_offset[r] .. tree_offset[r+1]-1, position 0 the root,
 * parent_pos[p] (relative to the root's first position) < p, pos_coord[p] the coordinate of the
 * edge to the parent or -1.  Candidates are in rank order: cand_root, the positions of the two
 * endpoints in that root's list, and cand_coord, the coordinate of the closing edge or -1.
 * ------------------------------------------------------------------------------------------- */

/* Parity of every position for S: par[p] = par[parent] ^ S[pos_coord[p]], par[root] = 0
 * (worker_thread::precompute_supportVec, include/work_per_thread.h:154-210, per position). */
int oracle_position_parity(int32_t F, const int64_t *tree_offset, const uint32_t *parent_pos,
                           const int32_t *pos_coord, int32_t W, const uint64_t *S, uint8_t *par,
                           int32_t num_threads);

/* One phase's list scan (mcb/mcb_cycle.cpp:254-285; mcb/mcb_cycle_deg_two_remove.cpp:329-358):
 * the first rank k with taken[k] == 0 (taken may be NULL) whose cycle is odd for S, i.e.
 * S[cand_coord[k]] ^ par[u] ^ par[v] = 1.  Returns that rank, -1 when there is none, -3 on bad
 * arguments, -5 when out of memory. */
int64_t oracle_phase_select(int32_t F, const int64_t *tree_offset, const uint32_t *parent_pos,
                            const int32_t *pos_coord, int64_t K, const uint32_t *cand_root,
                            const uint32_t *cand_pos_u, const uint32_t *cand_pos_v, const int32_t *cand_coord,
                            int32_t W, const uint64_t *S, const uint8_t *taken, int32_t num_threads);

/* cycle::get_cycle_vector (include/cycles.h:144-195) of rank k: the coordinates on the paths
 * root -> u and root -> v and of the closing edge, as W words in C. */
int oracle_cycle_vector(const int64_t *tree_offset, const uint32_t *parent_pos, const int32_t *pos_coord,
                        int64_t K, const uint32_t *cand_root, const uint32_t *cand_pos_u,
                        const uint32_t *cand_pos_v, const int32_t *cand_coord, int64_t k, int32_t W,
                        uint64_t *C);

/* Orthogonality of later support vectors to earlier cycle vectors (the de Pina invariant
 * <C_i, S_j> = 0 for i < j, mcb/mcb_cycle.cpp:296-301), tested for every i < j at once through 64
 * random combinations.  r[i] (i < N) is a random 64-bit word per phase; bit t of out[s] is
 * XOR over i < samp[s] with bit t of r[i] set of <C_i, S_samp[s]>, where C_i is the cycle vector
 * of rank selected[i].  samp must be ascending (each < N), S_samp holds their support vectors
 * (nsamp x W).  All-zero outputs mean every tested pair is orthogonal (a single nonzero product
 * leaves a given bit zero with probability 1/2, independently per bit). */
int oracle_orthogonality(const int64_t *tree_offset, const uint32_t *parent_pos, const int32_t *pos_coord,
                         int64_t K, const uint32_t *cand_root, const uint32_t *cand_pos_u,
                         const uint32_t *cand_pos_v, const int32_t *cand_coord, int32_t N,
                         const int32_t *selected, const uint64_t *r, int32_t nsamp, const int32_t *samp,
                         int32_t W, const uint64_t *S_samp, uint64_t *out);

#ifdef __cplusplus
}
#endif

#endif
