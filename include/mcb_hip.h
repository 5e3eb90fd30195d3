/*
 * mcb_hip.h - C ABI of the gfx950 device library (libmcb_hip.so).
 *
 * Two forms of the de Pina support-vector hot path live behind this boundary.
 *
 * 1. The reference-compatible form.  The five extern "C" entry points of gpu/common.cuh:18-27
 *    keep their names and signatures, and the C++ class gpu_struct (gpu/gpu_struct.cuh:15-84)
 *    becomes the opaque handle mcb_gpu with one function per method.  The trees, the coordinate
 *    map and the parity output keep the compressed_trees chunk layout of
 *    include/compressed_trees.h:16-26 (per chunk row: tree_rows u32[chunk*(n+1)],
 *    tree_cols u32[chunk*n], edge_offset i32[chunk*n], precompute_value u32[chunk*n]).
 *    A host driver written against the reference (mcb/mcb_cycle_deg_two_remove.cpp) can call it
 *    through a thin shim: a class gpu_struct whose methods forward to these functions.
 *
 * 2. The native form (mcb_solver_*).  Trees, candidates and every support vector stay resident in
 *    HBM; each phase runs (a) tree-path parity fused with (b) the minimum-rank odd-candidate
 *    search, the cycle-vector construction, and (c) the GF(2) support-vector update, without a
 *    host round trip.  Its inputs are the arrays of libmcb_host.so (include/mcb_host.h).
 *
 * Status-returning functions return MCB_OK (0) or a negative MCB_E_* code from mcb_host.h;
 * mcb_hip_last_error() gives the calling thread's last message.  The five reference names keep
 * the reference's behaviour on a runtime error: print and exit (include/utils.h:36-47).
 */
#ifndef MCB_HIP_H
#define MCB_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "mcb_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------
 * Reference entry points (gpu/common.cuh:18-27, gpu/common.cu:9-67)
 * ------------------------------------------------------------------------------------------- */

/* gpu/common.cu:9-12 - selects the current device and caches its properties. */
void init_cuda(void);

/* gpu/common.cu:14-30 - min(end - start, number of compute units). */
size_t configure_grid(int start, int end);

/* gpu/common.cu:32-40 - chunk*nodes uint32 of pinned host memory (not zeroed). */
unsigned *allocate_pinned_memory(int chunk, int nodes);

/* gpu/common.cu:42-46 */
void free_pinned_memory(unsigned *ptr);

/* gpu/common.cu:48-67 - largest chunk of roots whose trees fit in device memory next to the
 * coordinate map and the support vector, with nstream chunk slots. */
size_t calculate_chunk_size(size_t num_nodes, size_t num_edges, size_t size_vector, size_t nstream);

/* ---------------------------------------------------------------------------------------------
 * gpu_struct as an opaque handle (gpu/gpu_struct.cuh:15-84, gpu/gpu_struct.cu)
 * ------------------------------------------------------------------------------------------- */

typedef struct mcb_gpu mcb_gpu;

/* gpu_struct::gpu_struct (gpu_struct.cuh:44-59): num_edges = 2m', num_non_tree_edges = N,
 * size_vector = W, original_nodes = n', fvs_size = F, chunk_size roots per chunk, nstreams =
 * number of chunks (the reference passes ceil(F/chunk_size)).  Every chunk gets a device slot when
 * they fit in device memory; otherwise min(32, chunks) slots are allocated and trees must be
 * streamed per phase (multiple_transfer). */
int mcb_gpu_create(int num_edges, int num_non_tree_edges, int size_vector, int original_nodes, int fvs_size,
                   int chunk_size, int nstreams, mcb_gpu **out);

/* mcb_gpu_create with at most max_slots chunk slots on the device (0: as mcb_gpu_create).  With fewer
 * slots than chunks the trees stay on the host and mcb_gpu_process_shortest_path must be called with
 * multiple_transfer, which sends each chunk's trees to slot (chunk mod slots) before its parity pass
 * (gpu/gpu_struct.cu:132-156,175-186 stream the same way past 32 chunks).  No counterpart in the
 * reference's signatures: it lets a caller bound the device memory of the trees. */
int mcb_gpu_create_slots(int num_edges, int num_non_tree_edges, int size_vector, int original_nodes, int fvs_size,
                         int chunk_size, int nstreams, int max_slots, mcb_gpu **out);

/* gpu_struct::initialize_memory (gpu_struct.cu:77-102): uploads the coordinate map (num_edges
 * int32, -1 for global tree edges) and, when every chunk has a slot, every chunk's tree_cols and
 * edge_offset.  tree_rows is accepted for layout compatibility; the kernels only need the parent
 * position of each position, which must precede it (tree_cols[0] = 0xFFFFFFFF for the root). */
int mcb_gpu_initialize_memory(mcb_gpu *g, const int32_t *non_tree_edges_array, const uint32_t *const *tree_rows,
                              const uint32_t *const *tree_cols, const int32_t *const *edge_offset, int num_rows);

/* gpu_struct::copy_support_vector (gpu_struct.cu:104-114): S (size_vector uint64 words) to device.
 * *ms (optional) receives the elapsed time in milliseconds. */
int mcb_gpu_copy_support_vector(mcb_gpu *g, const uint64_t *elements, float *ms);

/* gpu_struct::process_shortest_path (gpu_struct.cu:169-193): for every chunk, parity of every
 * tree position for the current S (0/1 per position, the root 0), copied into
 * precompute_value[chunk].  With multiple_transfer the chunk's trees are sent first. */
int mcb_gpu_process_shortest_path(mcb_gpu *g, const uint32_t *const *tree_rows, const uint32_t *const *tree_cols,
                                  const int32_t *const *edge_offset, uint32_t *const *precompute_value, int num_rows,
                                  int multiple_transfer, float *ms);

/* gpu_struct::fetch (gpu_struct.cu:116-130): copies every resident chunk's parity to the host. */
int mcb_gpu_fetch(mcb_gpu *g, uint32_t *const *precompute_value, int num_rows, float *ms);

/* gpu_struct::calculate_memory (gpu_struct.cu:52-75), in MB. */
int mcb_gpu_calculate_memory(mcb_gpu *g, double *total_mb, double *static_mb, double *variable_mb);

/* gpu_struct::clear_memory (gpu_struct.cu:19-28) and destruction. */
void mcb_gpu_destroy(mcb_gpu *g);

/* ---------------------------------------------------------------------------------------------
 * Native device-resident phase loop
 * ------------------------------------------------------------------------------------------- */

typedef struct mcb_solver mcb_solver;

/* Device-ready problem (the arrays of include/mcb_host.h, all in host memory; copied at create). */
typedef struct {
    int32_t num_roots;               /* F */
    int32_t num_non_tree;            /* N */
    int64_t num_candidates;          /* K */
    int64_t num_positions;           /* P = tree_offset[F] */
    const int64_t *tree_offset;      /* [F+1] */
    const uint32_t *parent_pos;      /* [P] parent position within the root's layout, root 0xFFFFFFFF */
    const int32_t *pos_coord;        /* [P] coordinate of the edge to the parent or -1 */
    const uint32_t *cand_root;       /* [K] rank order */
    const uint32_t *cand_pos_u;      /* [K] */
    const uint32_t *cand_pos_v;      /* [K] */
    const int32_t *cand_coord;       /* [K] */
    const int64_t *root_cand_offset; /* [F+1] */
    const uint32_t *root_cand_rank;  /* [K] ascending within each root */
} mcb_solver_desc;

typedef struct {
    int32_t store_cycles;     /* 1: keep every phase's cycle vector (N*W words) for mcb_solver_cycle_vectors */
    int32_t waves_per_block;  /* blocked form's tree pass: waves per workgroup, 4, 8 or 16; 0 = the fastest
                                 of those whose LDS fits, timed on the first block at create */
    int32_t max_blocks;       /* 0 = automatic (resident blocks on every compute unit) */
    int32_t timing;           /* HIP events on the solver's stream around kernel launches: 0 none,
                                 1 every kernel, 2 only MCB_K_PARITY_SEARCH (the dominant kernel) and
                                 MCB_K_UPDATE (the support vector update) */
    int32_t block_phases;     /* phases resolved per pass over the trees: 0 = automatic (64),
                                 1 = one kernel per phase, 2..64 = blocked with that many */
    int32_t tree_cap;         /* blocked form's tree pass: largest tree whose path words stay in LDS, in
                                 positions (larger ones use global scratch); 0 = timed at create among
                                 as-large-as-fits, 8192 and 2048; -1 = as large as fits */
    int32_t merge_group;      /* blocked form's greedy basis: first-level bases merged per merge wave,
                                 2..64; 0 = automatic */
    int32_t timing_every;     /* with timing on, only every timing_every-th block of phases (one kernel per
                                 phase: every timing_every-th phase; batch: step) records events, counted
                                 from the last reset; 0 or 1 = every one.  An event pair costs the stream a
                                 few microseconds, so a sample keeps a timed loop close to an untimed one. */
    int32_t segment_resolve;  /* blocked tree pass, narrow trees in global scratch walked a segment per wave:
                                 0 = resolve only the segments' last positions in order and join the rest at
                                 the candidates' ends, where every segment's anchors allow it (the default);
                                 1 = resolve every segment in order.  Both give the same words. */
    int32_t dot_workgroups;   /* blocked form's support vector update: workgroup cap of the dot product
                                 launch; 0 = timed at create over the first blocks of phases among 512, 1024
                                 and 2048 (1024 without timing below 256 blocks of phases), 1..65536 = that
                                 cap.  Every cap gives the same results. */
    int32_t xor_workgroups;   /* the same for the XOR launch, timed among 1024, 2048 and 4096 (2048) */
} mcb_solver_options;

void mcb_solver_options_init(mcb_solver_options *opt);

int mcb_solver_create(const mcb_solver_desc *desc, const mcb_solver_options *opt, mcb_solver **out);

/* Sets every support vector back to its unit vector and the phase counter to 0. */
int mcb_solver_reset(mcb_solver *s);

/* Enqueues the next `phases` phases on the solver's stream (asynchronous). */
int mcb_solver_run(mcb_solver *s, int32_t phases);

/* The whole loop (phase 0 to N) as one HIP graph: captured at the first call, replayed after each
 * reset.  Needs phase 0 and timing off.  Same results as mcb_solver_run(s, N). */
int mcb_solver_run_graph(mcb_solver *s);

/* Waits for the stream; returns MCB_E_NOT_FOUND if a phase found no odd candidate. */
int mcb_solver_sync(mcb_solver *s);

/* Phases enqueued so far. */
int32_t mcb_solver_phase(const mcb_solver *s);

/* Rank of the candidate chosen in each of the first `count` phases. */
int mcb_solver_selected(mcb_solver *s, int32_t *out, int32_t count);

/* Support vectors as used in their phases, row-major N x W uint64 (row e = S_e). */
int mcb_solver_support_vectors(mcb_solver *s, uint64_t *out);

/* Support vectors S_j of the `count` phases listed in phases[] (each in 0..N-1), as used in their
 * phases, row-major count x W: the rows of mcb_solver_support_vectors without copying all N x W words
 * to the host.  Rows of phases not run yet hold their current state. */
int mcb_solver_support_vector_rows(mcb_solver *s, const int32_t *phases, int32_t count, uint64_t *out);

/* Cycle vectors of the first `count` phases, row-major count x W (needs store_cycles). */
int mcb_solver_cycle_vectors(mcb_solver *s, uint64_t *out, int32_t count);

/* Parity of every tree position for a given support vector S (W words), without searching:
 * out[p] = XOR of S bits along the path root -> position p (kernel (a) alone). */
int mcb_solver_parity(mcb_solver *s, const uint64_t *S, uint8_t *out);

/* mcb_solver_parity without device allocation: no counterpart in the reference.  For callers that
 * cannot allocate device memory per call (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  S and out are host arrays as in mcb_solver_parity; S and the parity bytes are staged in a
 * 64 MiB device array of the library (one per device, held until the process ends), so 8 * W rounded up
 * to 256 bytes plus the number of tree positions must fit in 64 MiB (MCB_E_INVALID otherwise).  Calls on
 * one device share that array: calls that overlap on one device race.  Runs on the current device, the
 * one the solver was created on. */
int mcb_solver_parity_noalloc(mcb_solver *s, const uint64_t *S, uint8_t *out);

/* Kernel ids for timing.  One kernel per phase: every launch of the phase kernel (update of the
 * previous phase, parity + search, cycle vector) counts under MCB_K_PARITY_SEARCH.  Blocked: per
 * block of phases, the pass over the trees (path words and candidate words) counts under
 * MCB_K_PARITY_SEARCH, the greedy basis and phase assignment under MCB_K_SELECT, the cycle and final
 * support vectors under MCB_K_CYCLE, the update of later support vectors under MCB_K_UPDATE and the
 * packing of the block's vectors under MCB_K_TRANSPOSE. */
enum {
    MCB_K_PARITY_SEARCH = 0,
    MCB_K_SELECT = 1,
    MCB_K_CYCLE = 2,
    MCB_K_UPDATE = 3,
    MCB_K_TRANSPOSE = 4,
    MCB_K_COUNT = 5
};

/* Accumulated event time (ms) and launch count of one kernel since the last reset (the kernels
 * that the timing option times). */
int mcb_solver_kernel_time(mcb_solver *s, int kernel, double *total_ms, int64_t *launches);

/* Device bytes the solver holds. */
int64_t mcb_solver_device_bytes(const mcb_solver *s);

/* Work done by the phase kernel since the last reset: tree positions streamed and candidates read
 * by the parity/search (roots whose best candidate cannot beat the phase's best are skipped), and
 * the algorithmic bytes of the support-vector update (8 per support word read for a dot product,
 * 16 per word XORed, 2 per column flag). */
int mcb_solver_work_counters(mcb_solver *s, int64_t *positions, int64_t *candidates, int64_t *update_bytes);

/* Launch shape of the dominant kernel (the phase kernel, or the blocked form's tree pass) and the
 * number of phases per block in use. */
int mcb_solver_launch_shape(const mcb_solver *s, int32_t *blocks, int32_t *waves_per_block, int32_t *lds_bytes,
                            int32_t *block_phases);

/* Workgroup caps of the support vector update's two launches in use (set or timed at create): no
 * counterpart in the reference, which has no such launches. */
int mcb_solver_update_grids(const mcb_solver *s, int32_t *dot_workgroups, int32_t *xor_workgroups);

void mcb_solver_destroy(mcb_solver *s);

/* ---------------------------------------------------------------------------------------------
 * Batched phase loop over many independent problems
 *
 * Replaces main.sh:1-30, which runs the MCB driver once per biconnected component written by
 * bicc/bicc_decomposition (one process, one device set-up and one phase loop each).  Here the
 * problems (normally the components of one graph, each preprocessed into an mcb_solver_desc) share
 * one device arena.  Step s resolves phases 64s .. 64s+63 of every problem that has them, with one
 * launch per kernel for all of them, so the number of launches follows the largest problem instead
 * of the sum over problems.  Each problem's selections are exactly those of its own mcb_solver.
 * ------------------------------------------------------------------------------------------- */

typedef struct mcb_batch mcb_batch;

/* descs[0..count-1] are validated like mcb_solver_create's; of opt, store_cycles, timing,
 * timing_every, waves_per_block and tree_cap are used (the plan is timed on step 0).  Fails with MCB_E_INVALID when a tree does
 * not fit the blocked form's LDS budget. */
int mcb_batch_create(const mcb_solver_desc *descs, int32_t count, const mcb_solver_options *opt, mcb_batch **out);

/* Back to the unit support vectors, step 0. */
int mcb_batch_reset(mcb_batch *b);

/* Steps in the whole loop (ceil of the largest phase count / 64), and the next step to run. */
int32_t mcb_batch_steps(const mcb_batch *b);
int32_t mcb_batch_step(const mcb_batch *b);

/* Enqueues the next `steps` steps (fewer at the end). */
int mcb_batch_run(mcb_batch *b, int32_t steps);

/* Every step as one HIP graph (captured at the first call, replayed after each reset); needs step 0
 * and timing off. */
int mcb_batch_run_graph(mcb_batch *b);

/* Waits; MCB_E_NOT_FOUND if some phase of some problem found no odd candidate. */
int mcb_batch_sync(mcb_batch *b);

/* Candidate rank (the problem's own numbering) chosen in phases 0..count-1 of problem `problem`
 * (index into descs); -1 for a phase not run yet. */
int mcb_batch_selected(mcb_batch *b, int32_t problem, int32_t *out, int32_t count);

/* Cycle vectors (count x W words of that problem) when created with store_cycles. */
int mcb_batch_cycle_vectors(mcb_batch *b, int32_t problem, uint64_t *out, int32_t count);

/* As mcb_solver_kernel_time, mcb_solver_device_bytes and mcb_solver_work_counters. */
int mcb_batch_kernel_time(mcb_batch *b, int kernel, double *total_ms, int64_t *launches);
int64_t mcb_batch_device_bytes(const mcb_batch *b);
int mcb_batch_work_counters(mcb_batch *b, int64_t *positions, int64_t *candidates, int64_t *update_bytes);

/* Launch shape of the tree pass and the number of steps. */
int mcb_batch_launch_shape(const mcb_batch *b, int32_t *blocks, int32_t *waves_per_block, int32_t *lds_bytes,
                           int32_t *steps);

void mcb_batch_destroy(mcb_batch *b);

/* ---------------------------------------------------------------------------------------------
 * Shortest-path trees of the FVS roots on the device.  Replaces the host's per-root searches
 * (include/Dijkstra.h:78-127 with the candidate filter of include/Dijkstra.h:205-254, run one root per
 * OpenMP thread at mcb/mcb_cycle_deg_two_remove.cpp:232-242 and include/work_per_thread.h:92-147) for
 * the pruned layout: every root's stopped search, with the reference heap's pop order, on one wave per
 * root.  Its results feed mcb_problem_finish_trees (include/mcb_host.h) of a problem built with
 * external_trees=1, which ranks the candidates and lays the trees out exactly as the host search would.
 * Errors: mcb_sptree_last_error().
 * ------------------------------------------------------------------------------------------- */

typedef struct mcb_sptree mcb_sptree;

typedef struct {
    int64_t slabs;          /* per-wave workspaces (waves of the widest stage) */
    int64_t launches;
    int64_t stage_roots[3]; /* roots finished in each stage: LDS heap of 1/8, 1/2, all of a CU's LDS */
    int64_t host_roots;     /* roots whose heap outgrew the last stage run: left to the host library */
    int64_t candidates;     /* candidate and layout entries handed over */
    int64_t kept;
    double device_ms;       /* all stages on the device, launches to the last status read */
} mcb_sptree_stats;

/* Uploads the reduced multigraph (the MCB_A_* arrays of the same names, include/mcb_host.h) and sizes
 * the per-wave workspaces (at most a quarter of the free device memory). */
int mcb_sptree_create(int32_t nodes, int32_t directed_edges, const uint32_t *row_offsets, const uint32_t *rows,
                      const uint32_t *cols, const int32_t *weights, const uint32_t *reverse, const int32_t *fvs_rank,
                      const uint32_t *roots, int32_t num_roots, mcb_sptree **out);

/* Searches every root, in up to `stages` (1..3) launches of growing LDS heap; a root whose heap outgrows
 * a stage is searched again in the next, unless that stage finished under an eighth of its roots (then
 * the rest are left to the host).  Blocks until the results are on the host. */
int mcb_sptree_run(mcb_sptree *t, int stages);

/* Waves per CU of each stage (count 1..3, each 1..16, not increasing; default 8, 2, 1): a stage's LDS heap
 * holds the CU's LDS share of one wave at 12 bytes per entry.  A count below 3 repeats the last value. */
int mcb_sptree_set_stages(mcb_sptree *t, const int32_t *waves_per_cu, int count);

/* The results of the last run in the form mcb_problem_finish_trees takes (valid until the next run or
 * mcb_sptree_destroy); roots left to the host have status != 0. */
int mcb_sptree_results(const mcb_sptree *t, mcb_root_trees *view);

int mcb_sptree_get_stats(const mcb_sptree *t, mcb_sptree_stats *stats);

void mcb_sptree_destroy(mcb_sptree *t);

const char *mcb_sptree_last_error(void);

/* ---------------------------------------------------------------------------------------------
 * Utilities
 * ------------------------------------------------------------------------------------------- */

/* Number of visible devices (0 when no GPU); never aborts. */
int mcb_hip_device_count(void);

/* Name and compute-unit count of the current device. */
int mcb_hip_device_info(char *name, int name_len, int *compute_units, int64_t *total_mem);

/* Copy kernel of `bytes` bytes between two device buffers, timed with events: the measured
 * bandwidth ceiling used beside the 8 TB/s HBM peak.  Returns GB/s in *gbps. */
int mcb_hip_copy_bandwidth(int64_t bytes, int iters, double *gbps);

const char *mcb_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MCB_HIP_H */
