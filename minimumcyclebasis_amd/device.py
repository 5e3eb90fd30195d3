"""Python front end of the gfx950 device library (``libmcb_hip.so``).

* ``Solver``    - the native device-resident phase loop over a ``host.Problem``.
* ``GpuStruct`` - the reference-compatible chunked parity interface (gpu_struct,
  gpu/gpu_struct.cuh:15-84) over numpy chunk arrays in the compressed_trees layout.
* ``SpTreeSearch`` - the roots' shortest-path trees searched on the device; ``problem_from_edges`` builds a
  problem with them (``trees="device"``) or on the host (``trees="host"``), with identical arrays.
* ``minimum_cycle_basis`` - end to end: preprocess, run every phase on the GPU.

There is no CPU fallback: without the HIP library or a GPU every call raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._hipsig import SolverDesc, SolverOptions, SpTreeStats
from .host import Problem
from .stats import RunStats, stats_of

KERNELS = ("parity_search", "select", "cycle", "update", "transpose")


def device_count() -> int:
    return int(_lib.hip().mcb_hip_device_count())


def device_info() -> dict:
    buf = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    mem = ctypes.c_int64()
    _lib.check_hip(_lib.hip().mcb_hip_device_info(buf, 256, ctypes.byref(cus), ctypes.byref(mem)))
    return {"name": buf.value.decode(), "compute_units": cus.value, "total_mem": mem.value}


def copy_bandwidth(nbytes: int = 1 << 30, iters: int = 20) -> float:
    """GB/s of a plain device copy kernel (read + write bytes)."""
    g = ctypes.c_double()
    _lib.check_hip(_lib.hip().mcb_hip_copy_bandwidth(int(nbytes), int(iters), ctypes.byref(g)))
    return g.value


def _desc_of(problem: Problem):
    """The solver descriptor of a preprocessed problem and the host arrays it points into (keep them alive)."""
    info = problem.info
    keep = [problem.array(n) for n in ("tree_offset", "parent_pos", "pos_coord", "cand_root", "cand_pos_u",
                                       "cand_pos_v", "cand_coord", "root_cand_offset", "root_cand_rank")]
    d = SolverDesc()
    d.num_roots = info.fvs_size
    d.num_non_tree = info.num_non_tree
    d.num_candidates = info.num_candidates
    d.num_positions = info.positions
    (d.tree_offset, d.parent_pos, d.pos_coord, d.cand_root, d.cand_pos_u, d.cand_pos_v, d.cand_coord,
     d.root_cand_offset, d.root_cand_rank) = [a.ctypes.data for a in keep]
    return d, keep


class Solver:
    """Device-resident de Pina phase loop for one preprocessed problem."""

    def __init__(self, problem: Problem, store_cycles: bool = False, timing: int | bool = False,
                 waves_per_block: int = 0, max_blocks: int = 0, block_phases: int = 0, tree_cap: int = 0,
                 merge_group: int = 0, timing_every: int = 0, segment_resolve: int = 0, dot_workgroups: int = 0,
                 xor_workgroups: int = 0):
        """timing: 0/False none, 1/True every kernel, 2 the tree pass and the update only (HIP events).
        timing_every: with timing on, only every timing_every-th block of phases records events (0, 1 all).
        block_phases: phases per pass over the trees (0 automatic, 1 one kernel per phase, up to 64).
        waves_per_block: waves per workgroup of the blocked tree pass (4, 8, 16; 0 timed at create).
        tree_cap: largest tree kept in LDS by the tree pass (0 timed at create, -1 as large as fits).
        merge_group: first-level bases merged per wave of the greedy basis (0 automatic, 2..64).
        segment_resolve: narrow trees in scratch: 0 rings first where the anchors allow, 1 every segment in
        order (the same words either way).
        dot_workgroups, xor_workgroups: workgroup caps of the support vector update's two launches (0 timed
        at create over the first blocks, 1..65536 that cap; the same results either way)."""
        info = problem.info
        if info.single_cycle:
            raise ValueError("a single-cycle graph has no phase loop")
        self.problem = problem
        self.N = info.num_non_tree
        self.W = info.words
        self.F = info.fvs_size
        self.K = info.num_candidates
        self.P = info.positions
        d, self._keep = _desc_of(problem)
        o = SolverOptions()
        _lib.hip().mcb_solver_options_init(ctypes.byref(o))
        o.store_cycles = int(bool(store_cycles))
        o.timing = int(timing) if isinstance(timing, int) and not isinstance(timing, bool) else int(bool(timing))
        o.waves_per_block = int(waves_per_block)
        o.max_blocks = int(max_blocks)
        o.block_phases = int(block_phases)
        o.tree_cap = int(tree_cap)
        o.merge_group = int(merge_group)
        o.segment_resolve = int(segment_resolve)
        o.timing_every = int(timing_every)
        o.dot_workgroups = int(dot_workgroups)
        o.xor_workgroups = int(xor_workgroups)
        h = ctypes.c_void_p()
        _lib.check_hip(_lib.hip().mcb_solver_create(ctypes.byref(d), ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self.store_cycles = bool(store_cycles)

    def reset(self) -> None:
        _lib.check_hip(_lib.hip().mcb_solver_reset(self._h))

    def run(self, phases: int | None = None) -> None:
        """Enqueue the next ``phases`` phases (all remaining when None)."""
        n = self.N if phases is None else int(phases)
        _lib.check_hip(_lib.hip().mcb_solver_run(self._h, n))

    def run_graph(self) -> None:
        """The whole loop from phase 0 as one HIP graph (captured once, replayed after each reset; timing off)."""
        _lib.check_hip(_lib.hip().mcb_solver_run_graph(self._h))

    def sync(self) -> None:
        _lib.check_hip(_lib.hip().mcb_solver_sync(self._h))

    @property
    def phase(self) -> int:
        return int(_lib.hip().mcb_solver_phase(self._h))

    def selected(self, count: int | None = None) -> np.ndarray:
        count = self.phase if count is None else int(count)
        out = np.zeros(count, dtype=np.int32)
        _lib.check_hip(_lib.hip().mcb_solver_selected(self._h, out.ctypes.data, count))
        return out

    def support_vectors(self) -> np.ndarray:
        out = np.zeros((self.N, self.W), dtype=np.uint64)
        _lib.check_hip(_lib.hip().mcb_solver_support_vectors(self._h, out.ctypes.data))
        return out

    def support_vector_rows(self, phases) -> np.ndarray:
        """Rows S_j (W words each) of the listed phases, without copying the whole N x W matrix."""
        ph = np.ascontiguousarray(phases, dtype=np.int32)
        out = np.zeros((ph.size, self.W), dtype=np.uint64)
        _lib.check_hip(_lib.hip().mcb_solver_support_vector_rows(self._h, ph.ctypes.data, int(ph.size),
                                                                  out.ctypes.data))
        return out

    def cycle_vectors(self, count: int | None = None) -> np.ndarray:
        count = self.phase if count is None else int(count)
        out = np.zeros((count, self.W), dtype=np.uint64)
        _lib.check_hip(_lib.hip().mcb_solver_cycle_vectors(self._h, out.ctypes.data, count))
        return out

    def parity(self, S: np.ndarray) -> np.ndarray:
        """Kernel (a) alone: parity of every position of every root layout for S."""
        S = np.ascontiguousarray(S, dtype=np.uint64)
        if S.size != self.W:
            raise ValueError("S must have W words")
        out = np.zeros(self.P, dtype=np.uint8)
        _lib.check_hip(_lib.hip().mcb_solver_parity(self._h, S.ctypes.data, out.ctypes.data))
        return out

    def parity_noalloc(self, S: np.ndarray) -> np.ndarray:
        """As ``parity``, staged in the library's fixed device scratch instead of allocating (at most 64 MiB:
        8 W bytes rounded up to 256 plus one byte per tree position)."""
        S = np.ascontiguousarray(S, dtype=np.uint64)
        if S.size != self.W:
            raise ValueError("S must have W words")
        out = np.zeros(self.P, dtype=np.uint8)
        _lib.check_hip(_lib.hip().mcb_solver_parity_noalloc(self._h, S.ctypes.data, out.ctypes.data))
        return out

    def kernel_time(self, kernel: str) -> tuple[float, int]:
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _lib.check_hip(_lib.hip().mcb_solver_kernel_time(self._h, KERNELS.index(kernel), ctypes.byref(ms),
                                                         ctypes.byref(n)))
        return ms.value, n.value

    @property
    def device_bytes(self) -> int:
        return int(_lib.hip().mcb_solver_device_bytes(self._h))

    def work_counters(self) -> tuple[int, int, int]:
        """(tree positions streamed, candidates read, support-update bytes) by the phase kernel since reset."""
        p = ctypes.c_int64()
        c = ctypes.c_int64()
        u = ctypes.c_int64()
        _lib.check_hip(_lib.hip().mcb_solver_work_counters(self._h, ctypes.byref(p), ctypes.byref(c), ctypes.byref(u)))
        return p.value, c.value, u.value

    def launch_shape(self) -> dict:
        v = [ctypes.c_int32() for _ in range(4)]
        _lib.check_hip(_lib.hip().mcb_solver_launch_shape(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("blocks", "waves_per_block", "lds_bytes", "block_phases"), (x.value for x in v)))

    def update_grids(self) -> dict:
        """Workgroup caps of the support vector update's launches in use (set or timed at create)."""
        d, x = ctypes.c_int32(), ctypes.c_int32()
        _lib.check_hip(_lib.hip().mcb_solver_update_grids(self._h, ctypes.byref(d), ctypes.byref(x)))
        return {"dot_workgroups": d.value, "xor_workgroups": x.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.hip().mcb_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Blocked phase loops of many independent problems in one device arena (``mcb_batch_*``): step s
    resolves phases 64s .. 64s+63 of every problem that has them, one launch per kernel for all of
    them.  Each problem's selections equal its own ``Solver``'s."""

    def __init__(self, problems, store_cycles: bool = False, timing: int | bool = False, waves_per_block: int = 0,
                 tree_cap: int = 0, merge_group: int = 0, timing_every: int = 0):
        problems = list(problems)
        if not problems:
            raise ValueError("no problems")
        for p in problems:
            if p.info.single_cycle:
                raise ValueError("a single-cycle graph has no phase loop")
        self.problems = problems
        self.N = [p.info.num_non_tree for p in problems]
        self.W = [p.info.words for p in problems]
        descs = (SolverDesc * len(problems))()
        self._keep = []
        for i, p in enumerate(problems):
            d, keep = _desc_of(p)
            descs[i] = d
            self._keep.append(keep)
        o = SolverOptions()
        _lib.hip().mcb_solver_options_init(ctypes.byref(o))
        o.store_cycles = int(bool(store_cycles))
        o.timing = int(timing) if isinstance(timing, int) and not isinstance(timing, bool) else int(bool(timing))
        o.waves_per_block = int(waves_per_block)
        o.tree_cap = int(tree_cap)
        o.merge_group = int(merge_group)
        o.timing_every = int(timing_every)
        h = ctypes.c_void_p()
        _lib.check_hip(_lib.hip().mcb_batch_create(ctypes.cast(descs, ctypes.c_void_p), len(problems),
                                                   ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self.store_cycles = bool(store_cycles)

    @property
    def steps(self) -> int:
        return int(_lib.hip().mcb_batch_steps(self._h))

    @property
    def step(self) -> int:
        return int(_lib.hip().mcb_batch_step(self._h))

    def reset(self) -> None:
        _lib.check_hip(_lib.hip().mcb_batch_reset(self._h))

    def run(self, steps: int | None = None) -> None:
        """Enqueue the next ``steps`` steps (all remaining when None)."""
        n = self.steps if steps is None else int(steps)
        _lib.check_hip(_lib.hip().mcb_batch_run(self._h, n))

    def run_graph(self) -> None:
        """Every step from step 0 as one HIP graph (captured once, replayed after each reset; timing off)."""
        _lib.check_hip(_lib.hip().mcb_batch_run_graph(self._h))

    def sync(self) -> None:
        _lib.check_hip(_lib.hip().mcb_batch_sync(self._h))

    def selected(self, problem: int, count: int | None = None) -> np.ndarray:
        count = self.N[problem] if count is None else int(count)
        out = np.zeros(count, dtype=np.int32)
        _lib.check_hip(_lib.hip().mcb_batch_selected(self._h, int(problem), out.ctypes.data, count))
        return out

    def cycle_vectors(self, problem: int, count: int | None = None) -> np.ndarray:
        count = self.N[problem] if count is None else int(count)
        out = np.zeros((count, self.W[problem]), dtype=np.uint64)
        _lib.check_hip(_lib.hip().mcb_batch_cycle_vectors(self._h, int(problem), out.ctypes.data, count))
        return out

    def kernel_time(self, kernel: str) -> tuple[float, int]:
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        _lib.check_hip(_lib.hip().mcb_batch_kernel_time(self._h, KERNELS.index(kernel), ctypes.byref(ms),
                                                        ctypes.byref(n)))
        return ms.value, n.value

    @property
    def device_bytes(self) -> int:
        return int(_lib.hip().mcb_batch_device_bytes(self._h))

    def work_counters(self) -> tuple[int, int, int]:
        """(tree positions streamed, candidates read, support-update bytes) since reset."""
        p = ctypes.c_int64()
        c = ctypes.c_int64()
        u = ctypes.c_int64()
        _lib.check_hip(_lib.hip().mcb_batch_work_counters(self._h, ctypes.byref(p), ctypes.byref(c), ctypes.byref(u)))
        return p.value, c.value, u.value

    def launch_shape(self) -> dict:
        v = [ctypes.c_int32() for _ in range(4)]
        _lib.check_hip(_lib.hip().mcb_batch_launch_shape(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("blocks", "waves_per_block", "lds_bytes", "steps"), (x.value for x in v)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.hip().mcb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuStruct:
    """gpu_struct (gpu/gpu_struct.cuh:15-84) over the compressed_trees chunk layout.

    ``tree_rows[i]``, ``tree_cols[i]``, ``edge_offset[i]`` and ``precompute_value[i]`` are the
    arrays of chunk row i (roots i*chunk .. i*chunk+chunk-1), each ``chunk * n`` entries
    (``chunk * (n + 1)`` for tree_rows), as include/compressed_trees.h:16-26 allocates them.
    """

    def __init__(self, num_edges: int, num_non_tree_edges: int, size_vector: int, original_nodes: int,
                 fvs_size: int, chunk_size: int, nstreams: int, max_slots: int = 0):
        """max_slots > 0 keeps at most that many chunk slots on the device; with fewer slots than chunks the
        trees are streamed every phase (process_shortest_path with multiple_transfer)."""
        h = ctypes.c_void_p()
        if max_slots:
            _lib.check_hip(_lib.hip().mcb_gpu_create_slots(num_edges, num_non_tree_edges, size_vector,
                                                           original_nodes, fvs_size, chunk_size, nstreams,
                                                           int(max_slots), ctypes.byref(h)))
        else:
            _lib.check_hip(_lib.hip().mcb_gpu_create(num_edges, num_non_tree_edges, size_vector, original_nodes,
                                                     fvs_size, chunk_size, nstreams, ctypes.byref(h)))
        self._h = h
        self.num_chunks = (fvs_size + chunk_size - 1) // chunk_size

    @staticmethod
    def _ptrs(arrays):
        arr = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
        return arr

    def initialize_memory(self, non_tree_edges: np.ndarray, tree_rows, tree_cols, edge_offset) -> None:
        nt = np.ascontiguousarray(non_tree_edges, dtype=np.int32)
        self._trees = (list(tree_rows), list(tree_cols), list(edge_offset))
        _lib.check_hip(_lib.hip().mcb_gpu_initialize_memory(self._h, nt.ctypes.data, self._ptrs(tree_rows),
                                                            self._ptrs(tree_cols), self._ptrs(edge_offset),
                                                            len(tree_cols)))

    def copy_support_vector(self, S: np.ndarray) -> float:
        S = np.ascontiguousarray(S, dtype=np.uint64)
        ms = ctypes.c_float()
        _lib.check_hip(_lib.hip().mcb_gpu_copy_support_vector(self._h, S.ctypes.data, ctypes.byref(ms)))
        return ms.value

    def process_shortest_path(self, precompute_value, multiple_transfer: bool = False) -> float:
        rows, cols, eo = self._trees
        ms = ctypes.c_float()
        _lib.check_hip(_lib.hip().mcb_gpu_process_shortest_path(
            self._h, self._ptrs(rows), self._ptrs(cols), self._ptrs(eo), self._ptrs(precompute_value),
            len(precompute_value), int(bool(multiple_transfer)), ctypes.byref(ms)))
        return ms.value

    def fetch(self, precompute_value) -> float:
        ms = ctypes.c_float()
        _lib.check_hip(_lib.hip().mcb_gpu_fetch(self._h, self._ptrs(precompute_value), len(precompute_value),
                                                ctypes.byref(ms)))
        return ms.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.hip().mcb_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpTreeSearch:
    """The stopped shortest-path searches of a problem's FVS roots on the device (mcb_sptree_*,
    include/mcb_hip.h); ``finish(problem)`` completes a problem built with ``Problem.for_external_trees``."""

    def __init__(self, problem: Problem):
        self._h = ctypes.c_void_p()
        info = problem.info
        keep = [problem.array(a) for a in ("row_offsets", "rows", "cols", "weights", "reverse", "fvs_rank", "roots")]
        rc = _lib.hip().mcb_sptree_create(info.nodes, info.directed_edges, *[a.ctypes.data for a in keep],
                                          info.fvs_size, ctypes.byref(self._h))
        self._check(rc)
        self.num_roots = info.fvs_size

    @staticmethod
    def _check(rc: int) -> int:
        if rc < 0:
            raise _lib.MCBError(rc, _lib.hip().mcb_sptree_last_error().decode(errors="replace"))
        return rc

    def set_stages(self, waves_per_cu) -> None:
        a = np.ascontiguousarray(waves_per_cu, dtype=np.int32)
        self._check(_lib.hip().mcb_sptree_set_stages(self._h, a.ctypes.data, int(a.size)))

    def run(self, stages: int = 3) -> None:
        self._check(_lib.hip().mcb_sptree_run(self._h, int(stages)))

    def view(self) -> _lib.RootTrees:
        v = _lib.RootTrees()
        self._check(_lib.hip().mcb_sptree_results(self._h, ctypes.byref(v)))
        return v

    def results(self) -> dict:
        """Copies of the per-root results (status, starts, counts, pools) as numpy arrays."""
        v = self.view()
        F = self.num_roots

        def arr(ptr, dt, count):
            if count == 0 or not ptr:
                return np.zeros(count, dtype=dt)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                         shape=(count,)).copy()

        out = {"status": arr(v.status, np.int32, F), "cand_start": arr(v.cand_start, np.int64, F),
               "cand_count": arr(v.cand_count, np.int32, F), "kept_start": arr(v.kept_start, np.int64, F),
               "kept_count": arr(v.kept_count, np.int32, F)}
        done = out["status"] == 0
        nc = int((out["cand_start"] + out["cand_count"])[done].max(initial=0))
        nk = int((out["kept_start"] + out["kept_count"])[done].max(initial=0))
        out["cand_edge"] = arr(v.cand_edge, np.uint32, nc)
        out["cand_weight"] = arr(v.cand_weight, np.int32, nc)
        out["kept"] = arr(v.kept, np.uint32, 4 * nk).reshape(nk, 4)
        return out

    def finish(self, problem: Problem, num_threads: int = 0) -> None:
        v = self.view()
        _lib.check_host(_lib.host().mcb_problem_finish_trees(problem.handle, ctypes.byref(v), int(num_threads)))
        problem._refresh()

    def stats(self) -> dict:
        st = SpTreeStats()
        self._check(_lib.hip().mcb_sptree_get_stats(self._h, ctypes.byref(st)))
        return {"slabs": st.slabs, "launches": st.launches, "stage_roots": list(st.stage_roots),
                "host_roots": st.host_roots, "candidates": st.candidates, "kept": st.kept,
                "device_ms": st.device_ms}

    def close(self) -> None:
        if self._h:
            _lib.hip().mcb_sptree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def problem_from_edges(n: int, u, v, w, trees: str = "device", reduce_degree_two: bool = True,
                       num_threads: int = 0, fvs_order=None, stages: int = 3, timings: dict | None = None) -> Problem:
    """A pruned-layout problem whose roots' shortest-path trees are searched on the device (trees="device")
    or by the host library (trees="host"); both give identical arrays.  ``timings`` (optional dict) receives
    the seconds of each part and the device search's statistics."""
    import time

    if trees == "host":
        t0 = time.perf_counter()
        p = Problem.from_edges(n, u, v, w, reduce_degree_two=reduce_degree_two, prune_trees=True,
                               num_threads=num_threads, fvs_order=fvs_order)
        if timings is not None:
            timings["host_s"] = time.perf_counter() - t0
        return p
    if trees != "device":
        raise ValueError("trees must be 'device' or 'host'")
    t0 = time.perf_counter()
    p = Problem.for_external_trees(n, u, v, w, reduce_degree_two=reduce_degree_two, num_threads=num_threads,
                                   fvs_order=fvs_order)
    t1 = time.perf_counter()
    if p.info.single_cycle:
        return p
    sp = SpTreeSearch(p)
    t2 = time.perf_counter()
    sp.run(stages)
    t3 = time.perf_counter()
    sp.finish(p, num_threads)
    t4 = time.perf_counter()
    if timings is not None:
        timings.update({"graph_fvs_coords_s": t1 - t0, "upload_s": t2 - t1, "device_search_s": t3 - t2,
                        "finish_s": t4 - t3, "search": sp.stats()})
    sp.close()
    return p


@dataclass
class MCBResult:
    total_weight: int            # exact (int64)
    total_weight_int32: int      # as the reference's int accumulator reports it
    num_cycles: int
    selected: np.ndarray         # candidate rank chosen in each phase
    cycles: tuple                # (root vertex, u, v, weight) arrays in phase order, reduced-graph ids
    stats: RunStats | None = None  # the reference's stats CSV row (minimumcyclebasis_amd.stats)


def minimum_cycle_basis(source, reduce_degree_two: bool = True, prune_trees: bool = True, num_threads: int = 0,
                        fvs_order=None, block_phases: int = 0, fvs: str = "plain", trees: str = "device") -> MCBResult:
    """Minimum cycle basis weight and cycles of a graph given as a .mtx path or (n, u, v, w).  The result's
    ``stats`` is the run's row for the reference's stats CSV (``stats.append_stats_csv``).

    fvs="learned" orders the FVS pass by the learned vertex scorer (``learned_fvs``), falling back to the
    reference order unless that gives a smaller verified FVS; the weight is the same either way.
    trees="device" searches the roots' shortest-path trees on the GPU (pruned layout, edge-list input),
    "host" in the host library; the problem is identical either way."""
    import time

    t0 = time.perf_counter()
    if fvs not in ("plain", "learned"):
        raise ValueError("fvs must be 'plain' or 'learned'")
    if fvs == "learned" and not isinstance(source, Problem) and fvs_order is None:
        from . import generators
        from .learned_fvs import learned_order

        if isinstance(source, (str, bytes)) or hasattr(source, "__fspath__"):
            source = generators.read_mtx(source)
        fvs_order = learned_order(*source, reduce_degree_two=reduce_degree_two).order
    if isinstance(source, Problem):
        prob = source
    elif isinstance(source, (str, bytes)) or hasattr(source, "__fspath__"):
        prob = Problem.from_mtx(source, reduce_degree_two=reduce_degree_two, prune_trees=prune_trees,
                                num_threads=num_threads, fvs_order=fvs_order)
    elif prune_trees and trees == "device":
        n, u, v, w = source
        prob = problem_from_edges(n, u, v, w, trees="device", reduce_degree_two=reduce_degree_two,
                                  num_threads=num_threads, fvs_order=fvs_order)
    else:
        n, u, v, w = source
        prob = Problem.from_edges(n, u, v, w, reduce_degree_two=reduce_degree_two, prune_trees=prune_trees,
                                  num_threads=num_threads, fvs_order=fvs_order)
    info = prob.info
    t_host = time.perf_counter() - t0
    if info.single_cycle:
        tw = int(info.input_total_weight)
        return MCBResult(tw, int(np.int32(np.int64(tw) & 0xFFFFFFFF).astype(np.int32)) if tw else 0, 1,
                         np.zeros(1, dtype=np.int32), (), stats_of(prob, tw, construction_trees_s=t_host))
    t1 = time.perf_counter()
    s = Solver(prob, timing=2, block_phases=block_phases)
    s.run()
    s.sync()
    t_loop = time.perf_counter() - t1
    sel = s.selected()
    dev_ms = sum(s.kernel_time(k)[0] for k in KERNELS)
    mem_mb = s.device_bytes / 2**20
    s.close()
    return _result_of(prob, sel, t_host, t_loop, dev_ms, mem_mb)


def _result_of(prob: Problem, sel: np.ndarray, t_host: float, t_loop: float, dev_ms: float,
               mem_mb: float) -> MCBResult:
    r, a, b, w = prob.candidate_table()
    ww = w[sel].astype(np.int64)
    total = int(ww.sum())
    st = stats_of(prob, total, construction_trees_s=t_host, inspection_s=t_loop, precompute_s=dev_ms / 1000.0,
                  total_memory_mb=mem_mb, static_memory_mb=mem_mb)
    return MCBResult(total, int(np.int64(total).astype(np.int32)), int(sel.size), sel,
                     (r[sel], a[sel], b[sel], w[sel]), st)


@dataclass
class ComponentsResult:
    total_weight: int            # exact (int64), sum over the components
    num_cycles: int
    components: list             # (components.Component, MCBResult) per biconnected component with a cycle


def minimum_cycle_basis_components(source, reduce_degree_two: bool = True, prune_trees: bool = True,
                                   num_threads: int = 0, block_phases: int = 0,
                                   batched: bool = False) -> ComponentsResult:
    """Minimum cycle basis of a graph solved one biconnected component at a time, in process (the reference runs
    main.sh: bicc decomposition, relabelling, one MCB run per component).  ``source`` is a .mtx path or
    (n, u, v, w).  Each component's result carries its stats row (``MCBResult.stats``).

    batched=False runs one ``Solver`` per component, one after the other.  batched=True runs every component's
    phase loop together in one ``Batch`` (same selections); each row's main-loop time is then the shared batch's."""
    import time

    from . import generators
    from .components import split_components

    if isinstance(source, (str, bytes)) or hasattr(source, "__fspath__"):
        n, u, v, w = generators.read_mtx(source)
    else:
        n, u, v, w = source
    comps = split_components(n, u, v, w)
    if not batched:
        parts = []
        for comp in comps:
            res = minimum_cycle_basis((comp.n, comp.u, comp.v, comp.w), reduce_degree_two=reduce_degree_two,
                                      prune_trees=prune_trees, num_threads=num_threads, block_phases=block_phases)
            parts.append((comp, res))
        return ComponentsResult(sum(r.total_weight for _, r in parts), sum(r.num_cycles for _, r in parts), parts)
    probs, t_host = [], []
    for comp in comps:
        t0 = time.perf_counter()
        probs.append(Problem.from_edges(comp.n, comp.u, comp.v, comp.w, reduce_degree_two=reduce_degree_two,
                                        prune_trees=prune_trees, num_threads=num_threads))
        t_host.append(time.perf_counter() - t0)
    loop = [i for i, p in enumerate(probs) if not p.info.single_cycle]
    results = {}
    for i, p in enumerate(probs):
        if p.info.single_cycle:
            results[i] = minimum_cycle_basis(p)
            results[i].stats.construction_trees_s = t_host[i]
    if loop:
        t1 = time.perf_counter()
        bt = Batch([probs[i] for i in loop], timing=2)
        bt.run()
        bt.sync()
        t_loop = time.perf_counter() - t1
        dev_ms = sum(bt.kernel_time(k)[0] for k in KERNELS)
        mem_mb = bt.device_bytes / 2**20
        for j, i in enumerate(loop):
            results[i] = _result_of(probs[i], bt.selected(j), t_host[i], t_loop, dev_ms, mem_mb)
        bt.close()
    parts = [(comps[i], results[i]) for i in range(len(comps))]
    return ComponentsResult(sum(r.total_weight for _, r in parts), sum(r.num_cycles for _, r in parts), parts)
