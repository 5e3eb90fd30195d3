"""ctypes signatures of libmcb_hip.so (include/mcb_hip.h)."""
from __future__ import annotations

import ctypes

P = ctypes.c_void_p
i32, i64, u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64


class SolverDesc(ctypes.Structure):
    _fields_ = [
        ("num_roots", i32),
        ("num_non_tree", i32),
        ("num_candidates", i64),
        ("num_positions", i64),
        ("tree_offset", P),
        ("parent_pos", P),
        ("pos_coord", P),
        ("cand_root", P),
        ("cand_pos_u", P),
        ("cand_pos_v", P),
        ("cand_coord", P),
        ("root_cand_offset", P),
        ("root_cand_rank", P),
    ]


class SolverOptions(ctypes.Structure):
    _fields_ = [
        ("store_cycles", i32),
        ("waves_per_block", i32),
        ("max_blocks", i32),
        ("timing", i32),
        ("block_phases", i32),
        ("tree_cap", i32),
        ("merge_group", i32),
        ("timing_every", i32),
        ("segment_resolve", i32),
        ("dot_workgroups", i32),
        ("xor_workgroups", i32),
    ]


class SpTreeStats(ctypes.Structure):
    _fields_ = [
        ("slabs", i64),
        ("launches", i64),
        ("stage_roots", i64 * 3),
        ("host_roots", i64),
        ("candidates", i64),
        ("kept", i64),
        ("device_ms", ctypes.c_double),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "init_cuda": (None, []),
    "configure_grid": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]),
    "allocate_pinned_memory": (P, [ctypes.c_int, ctypes.c_int]),
    "free_pinned_memory": (None, [P]),
    "calculate_chunk_size": (ctypes.c_size_t, [ctypes.c_size_t] * 4),
    "mcb_gpu_create": (ctypes.c_int, [ctypes.c_int] * 7 + [ctypes.POINTER(P)]),
    "mcb_gpu_create_slots": (ctypes.c_int, [ctypes.c_int] * 8 + [ctypes.POINTER(P)]),
    "mcb_gpu_initialize_memory": (ctypes.c_int, [P, P, P, P, P, ctypes.c_int]),
    "mcb_gpu_copy_support_vector": (ctypes.c_int, [P, P, ctypes.POINTER(ctypes.c_float)]),
    "mcb_gpu_process_shortest_path": (ctypes.c_int, [P, P, P, P, P, ctypes.c_int, ctypes.c_int,
                                                     ctypes.POINTER(ctypes.c_float)]),
    "mcb_gpu_fetch": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]),
    "mcb_gpu_calculate_memory": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                ctypes.POINTER(ctypes.c_double)]),
    "mcb_gpu_destroy": (None, [P]),
    "mcb_solver_options_init": (None, [ctypes.POINTER(SolverOptions)]),
    "mcb_solver_create": (ctypes.c_int, [ctypes.POINTER(SolverDesc), ctypes.POINTER(SolverOptions), ctypes.POINTER(P)]),
    "mcb_solver_reset": (ctypes.c_int, [P]),
    "mcb_solver_run": (ctypes.c_int, [P, i32]),
    "mcb_solver_sync": (ctypes.c_int, [P]),
    "mcb_solver_run_graph": (ctypes.c_int, [P]),
    "mcb_solver_phase": (i32, [P]),
    "mcb_solver_selected": (ctypes.c_int, [P, P, i32]),
    "mcb_solver_support_vectors": (ctypes.c_int, [P, P]),
    "mcb_solver_support_vector_rows": (ctypes.c_int, [P, P, i32, P]),
    "mcb_solver_cycle_vectors": (ctypes.c_int, [P, P, i32]),
    "mcb_solver_parity": (ctypes.c_int, [P, P, P]),
    "mcb_solver_parity_noalloc": (ctypes.c_int, [P, P, P]),
    "mcb_solver_kernel_time": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]),
    "mcb_solver_device_bytes": (i64, [P]),
    "mcb_solver_work_counters": (ctypes.c_int, [P, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)]),
    "mcb_solver_launch_shape": (ctypes.c_int, [P, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                               ctypes.POINTER(i32)]),
    "mcb_solver_update_grids": (ctypes.c_int, [P, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
    "mcb_solver_destroy": (None, [P]),
    "mcb_batch_create": (ctypes.c_int, [P, i32, ctypes.POINTER(SolverOptions), ctypes.POINTER(P)]),
    "mcb_batch_reset": (ctypes.c_int, [P]),
    "mcb_batch_steps": (i32, [P]),
    "mcb_batch_step": (i32, [P]),
    "mcb_batch_run": (ctypes.c_int, [P, i32]),
    "mcb_batch_sync": (ctypes.c_int, [P]),
    "mcb_batch_run_graph": (ctypes.c_int, [P]),
    "mcb_batch_selected": (ctypes.c_int, [P, i32, P, i32]),
    "mcb_batch_cycle_vectors": (ctypes.c_int, [P, i32, P, i32]),
    "mcb_batch_kernel_time": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]),
    "mcb_batch_device_bytes": (i64, [P]),
    "mcb_batch_work_counters": (ctypes.c_int, [P, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)]),
    "mcb_batch_launch_shape": (ctypes.c_int, [P, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                              ctypes.POINTER(i32)]),
    "mcb_batch_destroy": (None, [P]),
    "mcb_mlp_create": (ctypes.c_int, [ctypes.POINTER(P)]),
    "mcb_mlp_set_params": (ctypes.c_int, [P, P]),
    "mcb_mlp_get_params": (ctypes.c_int, [P, P]),
    "mcb_mlp_forward": (ctypes.c_int, [P, P, i64, P]),
    "mcb_mlp_gradient": (ctypes.c_int, [P, P, P, i64, P, ctypes.POINTER(ctypes.c_float)]),
    "mcb_mlp_set_data": (ctypes.c_int, [P, P, P, i64]),
    "mcb_mlp_train_epoch": (ctypes.c_int, [P, P, i64, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "mcb_mlp_destroy": (None, [P]),
    "mcb_mlp_last_error": (ctypes.c_char_p, []),
    "mcb_sptree_create": (ctypes.c_int, [i32, i32, P, P, P, P, P, P, P, i32, ctypes.POINTER(P)]),
    "mcb_sptree_run": (ctypes.c_int, [P, ctypes.c_int]),
    "mcb_sptree_set_stages": (ctypes.c_int, [P, P, ctypes.c_int]),
    "mcb_sptree_results": (ctypes.c_int, [P, P]),
    "mcb_sptree_get_stats": (ctypes.c_int, [P, ctypes.POINTER(SpTreeStats)]),
    "mcb_sptree_destroy": (None, [P]),
    "mcb_sptree_last_error": (ctypes.c_char_p, []),
    "mcb_hip_device_count": (ctypes.c_int, []),
    "mcb_hip_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(i64)]),
    "mcb_hip_copy_bandwidth": (ctypes.c_int, [i64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "mcb_hip_last_error": (ctypes.c_char_p, []),
}


def declare(lib: ctypes.CDLL) -> None:
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
