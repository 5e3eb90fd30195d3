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

#ifdef __cplusplus
}
#endif

#endif
