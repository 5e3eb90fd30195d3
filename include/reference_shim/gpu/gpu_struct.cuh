/*
 * gpu/gpu_struct.cuh for building the reference's GPU driver with g++ against libmcb_hip.so: the
 * class gpu_struct of the reference's gpu/gpu_struct.cuh:15-84 (same name, constructor and the
 * methods the driver calls), each method forwarding to the mcb_gpu handle of mcb_hip.h.
 *
 * Behaviour kept: every device error prints and exits, as CudaError does (include/utils.h:36-47);
 * process_shortest_path leaves precompute_value[chunk][pos] = parity of position pos
 * (gpu/gpu_struct.cu:158-167); the times are milliseconds of device events.
 * One difference: the reference uploads the coordinate map only in initialize_memory, which its
 * driver skips when it streams the trees (mcb_cycle_deg_two_remove.cpp:296-297), so its kernels then
 * read an uninitialised map; here the first process_shortest_path uploads the map when
 * initialize_memory has not run.
 */
#ifndef __H_GPU_STRUCT
#define __H_GPU_STRUCT

#include <cstdio>
#include <cstdlib>

#include "bit_vector.h"
#include "gpu_task.h"
#include "mcb_hip.h"
#include "stats.h"

struct gpu_struct {
    mcb_gpu *handle;
    bool initialized;
    int num_chunks;
    stats *info;

    gpu_struct(int num_edges, int num_non_tree_edges, int size_vector, int original_nodes, int fvs_size,
               int chunk_size, int nstreams, stats *info)
        : handle(NULL), initialized(false), num_chunks(nstreams), info(info) {
        check(mcb_gpu_create(num_edges, num_non_tree_edges, size_vector, original_nodes, fvs_size, chunk_size,
                             nstreams, &handle));
    }

    ~gpu_struct() { clear_memory(); }

    void calculate_memory() {
        double total = 0, fixed = 0, per_chunk = 0;
        check(mcb_gpu_calculate_memory(handle, &total, &fixed, &per_chunk));
        std::printf("Device memory %.2f MB (static %.2f MB, chunk slots %.2f MB)\n", total, fixed, per_chunk);
    }

    void initialize_memory(gpu_task *t) {
        compressed_trees *ht = t->host_tree;
        check(mcb_gpu_initialize_memory(handle, t->non_tree_edges_array, ht->tree_rows, ht->tree_cols,
                                        ht->edge_offset, ht->num_rows));
        initialized = true;
    }

    float copy_support_vector(bit_vector *v) {
        float ms = 0;
        check(mcb_gpu_copy_support_vector(handle, reinterpret_cast<const uint64_t *>(v->elements), &ms));
        return ms;
    }

    float process_shortest_path(gpu_task *t, bool multiple_transfer) {
        if (!initialized) initialize_memory(t);
        compressed_trees *ht = t->host_tree;
        float ms = 0;
        check(mcb_gpu_process_shortest_path(handle, ht->tree_rows, ht->tree_cols, ht->edge_offset,
                                            ht->precompute_value, ht->num_rows, multiple_transfer ? 1 : 0, &ms));
        return ms;
    }

    float fetch(gpu_task *t) {
        float ms = 0;
        check(mcb_gpu_fetch(handle, t->host_tree->precompute_value, t->host_tree->num_rows, &ms));
        return ms;
    }

    void clear_memory() {
        if (handle) mcb_gpu_destroy(handle);
        handle = NULL;
    }

  private:
    static void check(int rc) {
        if (rc) {
            std::fprintf(stderr, "Error : %s\n", mcb_hip_last_error());
            std::exit(1);
        }
    }
};

#endif
