"""Python view of a preprocessed MCB problem (``libmcb_host.so``).

``Problem`` owns one native ``mcb_problem``: the reduced multigraph, the FVS roots, the
cycle-space coordinates, the per-root shortest-path tree layouts and the ranked candidate
cycles.  It mirrors the host half of the reference driver
(mcb/mcb_cycle_deg_two_remove.cpp:68-267, mcb/mcb_cycle.cpp:76-203).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

# Array ids (include/mcb_host.h) with their numpy dtypes.
ARRAYS = {
    "rows": (0, np.uint32),
    "cols": (1, np.uint32),
    "weights": (2, np.int32),
    "reverse": (3, np.uint32),
    "row_offsets": (4, np.uint32),
    "fvs_rank": (5, np.int32),
    "roots": (6, np.uint32),
    "nt_edges": (7, np.uint32),
    "nt_map": (8, np.int32),
    "cand_root": (9, np.uint32),
    "cand_edge": (10, np.uint32),
    "cand_weight": (11, np.int32),
    "cand_pos_u": (12, np.uint32),
    "cand_pos_v": (13, np.uint32),
    "cand_coord": (14, np.int32),
    "tree_offset": (15, np.int64),
    "parent_pos": (16, np.uint32),
    "pos_coord": (17, np.int32),
    "root_cand_offset": (18, np.int64),
    "root_cand_rank": (19, np.uint32),
    "vertex_parent": (20, np.int32),
    "pos_vertex": (21, np.uint32),
}


@dataclass
class Info:
    input_nodes: int
    input_edges: int
    single_cycle: bool
    nodes_removed: int
    source_vertex: int
    nodes: int
    directed_edges: int
    num_non_tree: int
    words: int
    fvs_size: int
    num_candidates: int
    positions: int
    max_positions: int
    max_depth: int
    input_total_weight: int


def _options(reduce_degree_two: bool, prune_trees: bool, keep_vertex_trees: bool, num_threads: int,
             fvs_order=None, stop_after_fvs: bool = False, external_trees: bool = False):
    opt = _lib.BuildOptions()
    _lib.host().mcb_build_options_init(ctypes.byref(opt))
    opt.stop_after_fvs = int(bool(stop_after_fvs))
    opt.external_trees = int(bool(external_trees))
    opt.reduce_degree_two = int(bool(reduce_degree_two))
    opt.prune_trees = int(bool(prune_trees))
    opt.keep_vertex_trees = int(bool(keep_vertex_trees))
    opt.num_threads = int(num_threads)
    keep = None
    if fvs_order is not None:
        keep = np.ascontiguousarray(fvs_order, dtype=np.int32)
        opt.fvs_order = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        opt.fvs_order_len = int(keep.size)
    return opt, keep


class Problem:
    """A preprocessed graph ready for the phase loop.

    Build from a Matrix-Market file (``Problem.from_mtx``) or an edge list
    (``Problem.from_edges``, 0-based endpoints, each undirected edge once).
    """

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle
        info = _lib.ProblemInfo()
        _lib.check_host(_lib.host().mcb_problem_get_info(self._h, ctypes.byref(info)))
        self.info = Info(**{f: getattr(info, f) for f, _ in _lib.ProblemInfo._fields_})
        self.info.single_cycle = bool(self.info.single_cycle)
        self._cache: dict = {}

    @classmethod
    def from_mtx(cls, path: str, reduce_degree_two: bool = True, prune_trees: bool = False,
                 keep_vertex_trees: bool = False, num_threads: int = 0, fvs_order=None) -> "Problem":
        opt, _keep = _options(reduce_degree_two, prune_trees, keep_vertex_trees, num_threads, fvs_order)
        h = ctypes.c_void_p()
        _lib.check_host(_lib.host().mcb_problem_build_mtx(str(path).encode(), ctypes.byref(opt), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, n: int, u, v, w, reduce_degree_two: bool = True, prune_trees: bool = False,
                   keep_vertex_trees: bool = False, num_threads: int = 0, fvs_order=None) -> "Problem":
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.int32)
        if not (u.shape == v.shape == w.shape):
            raise ValueError("u, v, w must have the same length")
        opt, _keep = _options(reduce_degree_two, prune_trees, keep_vertex_trees, num_threads, fvs_order)
        h = ctypes.c_void_p()
        _lib.check_host(_lib.host().mcb_problem_build_edges(
            int(n), int(u.size), u.ctypes.data, v.ctypes.data, w.ctypes.data, ctypes.byref(opt), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def for_external_trees(cls, n: int, u, v, w, reduce_degree_two: bool = True, num_threads: int = 0,
                           fvs_order=None) -> "Problem":
        """The reduced graph, its FVS and coordinates, with the pruned shortest-path trees left to
        ``finish_trees`` (fed by the device search, ``device.SpTreeSearch``, or by ``root_search``)."""
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.int32)
        if not (u.shape == v.shape == w.shape):
            raise ValueError("u, v, w must have the same length")
        opt, _keep = _options(reduce_degree_two, True, False, num_threads, fvs_order, external_trees=True)
        h = ctypes.c_void_p()
        _lib.check_host(_lib.host().mcb_problem_build_edges(
            int(n), int(u.size), u.ctypes.data, v.ctypes.data, w.ctypes.data, ctypes.byref(opt), ctypes.byref(h)))
        return cls(h)

    def root_search(self, r: int):
        """The host's stopped search of root index ``r`` in the form the device search returns:
        (accepted candidate edges, their weights, kept vertices as rows of (vertex, parent edge, level,
        finalisation index))."""
        nc = ctypes.c_int32()
        nk = ctypes.c_int32()
        host = _lib.host()
        rc = host.mcb_problem_root_search(self._h, int(r), 0, None, None, 0, None, ctypes.byref(nc), ctypes.byref(nk))
        if rc not in (0, -3):
            _lib.check_host(rc)
        ce = np.zeros(nc.value, dtype=np.uint32)
        cw = np.zeros(nc.value, dtype=np.int32)
        kept = np.zeros((nk.value, 4), dtype=np.uint32)
        _lib.check_host(host.mcb_problem_root_search(self._h, int(r), ce.size, ce.ctypes.data, cw.ctypes.data,
                                                     kept.shape[0], kept.ctypes.data, ctypes.byref(nc),
                                                     ctypes.byref(nk)))
        return ce, cw, kept

    def finish_trees(self, status, cand_start, cand_count, cand_edge, cand_weight, kept_start, kept_count, kept,
                     num_threads: int = 0) -> None:
        """Completes a problem built by ``for_external_trees`` from per-root search results (roots whose
        status is nonzero are searched by the host library)."""
        arrs = [np.ascontiguousarray(status, dtype=np.int32), np.ascontiguousarray(cand_start, dtype=np.int64),
                np.ascontiguousarray(cand_count, dtype=np.int32), np.ascontiguousarray(cand_edge, dtype=np.uint32),
                np.ascontiguousarray(cand_weight, dtype=np.int32), np.ascontiguousarray(kept_start, dtype=np.int64),
                np.ascontiguousarray(kept_count, dtype=np.int32), np.ascontiguousarray(kept, dtype=np.uint32)]
        F = self.info.fvs_size
        if any(a.size != F for a in (arrs[0], arrs[1], arrs[2], arrs[5], arrs[6])):
            raise ValueError("per-root arrays must have one entry per root")
        t = _lib.RootTrees(*[a.ctypes.data for a in arrs])
        _lib.check_host(_lib.host().mcb_problem_finish_trees(self._h, ctypes.byref(t), int(num_threads)))
        self._refresh()

    def _refresh(self) -> None:
        info = _lib.ProblemInfo()
        _lib.check_host(_lib.host().mcb_problem_get_info(self._h, ctypes.byref(info)))
        self.info = Info(**{f: getattr(info, f) for f, _ in _lib.ProblemInfo._fields_})
        self.info.single_cycle = bool(self.info.single_cycle)
        self._cache.clear()

    @classmethod
    def reduced_graph(cls, n: int, u, v, w, reduce_degree_two: bool = True) -> "Problem":
        """Only the reduced multigraph and its reference-order FVS (rows, cols, weights, row offsets,
        fvs rank, roots; info through fvs_size and num_non_tree): cheap, no shortest-path trees."""
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.int32)
        opt, _keep = _options(reduce_degree_two, False, False, 1, None, stop_after_fvs=True)
        h = ctypes.c_void_p()
        _lib.check_host(_lib.host().mcb_problem_build_edges(
            int(n), int(u.size), u.ctypes.data, v.ctypes.data, w.ctypes.data, ctypes.byref(opt), ctypes.byref(h)))
        return cls(h)

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def array(self, name: str) -> np.ndarray:
        """A read-only numpy view of an internal array (valid while this object lives)."""
        if name in self._cache:
            return self._cache[name]
        which, dtype = ARRAYS[name]
        ptr = ctypes.c_void_p()
        cnt = ctypes.c_int64()
        _lib.check_host(_lib.host().mcb_problem_array(self._h, which, ctypes.byref(ptr), ctypes.byref(cnt)))
        n = int(cnt.value)
        if n == 0 or not ptr.value:
            arr = np.zeros(0, dtype=dtype)
        else:
            buf = (ctypes.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr.value)
            arr = np.frombuffer(buf, dtype=dtype, count=n)
        arr.flags.writeable = False
        self._cache[name] = arr
        return arr

    def reference_layout(self, r: int):
        """(rows, cols, edge_offset, nodes_index) of root index ``r`` in the reference's own
        BFS-level layout (include/Dijkstra.h:151-203)."""
        n = self.info.nodes
        rows = np.zeros(n + 1, dtype=np.uint32)
        cols = np.zeros(n, dtype=np.uint32)
        eo = np.zeros(n, dtype=np.int32)
        ni = np.zeros(n, dtype=np.uint32)
        _lib.check_host(_lib.host().mcb_problem_reference_layout(
            self._h, int(r), rows.ctypes.data, cols.ctypes.data, eo.ctypes.data, ni.ctypes.data))
        return rows, cols, eo, ni

    def candidate_table(self):
        """Candidates in rank order: (root vertex, u, v, weight) with 0-based reduced-graph ids."""
        roots = self.array("roots")
        e = self.array("cand_edge")
        return (roots[self.array("cand_root")], self.array("rows")[e], self.array("cols")[e],
                self.array("cand_weight"))

    def close(self) -> None:
        if self._h:
            self._cache.clear()
            _lib.host().mcb_problem_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fvs_reverse_delete(row_offsets, cols, order=None) -> np.ndarray:
    """Reverse-delete FVS of a CSR multigraph; returns the root rank per vertex (-1 = not in F)."""
    ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
    c = np.ascontiguousarray(cols, dtype=np.uint32)
    n = ro.size - 1
    out = np.zeros(n, dtype=np.int32)
    o = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
    _lib.check_host(_lib.host().mcb_fvs_reverse_delete(
        n, int(c.size), ro.ctypes.data, c.ctypes.data, None if o is None else o.ctypes.data, out.ctypes.data))
    return out


def induced_has_cycle(row_offsets, cols, in_set) -> bool:
    ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
    c = np.ascontiguousarray(cols, dtype=np.uint32)
    s = np.ascontiguousarray(in_set, dtype=np.uint8)
    return bool(_lib.check_host(_lib.host().mcb_induced_has_cycle(ro.size - 1, ro.ctypes.data, c.ctypes.data,
                                                                  s.ctypes.data)))


def biconnected_components(n: int, u, v) -> np.ndarray:
    """Component id per edge (simple graph, 0-based endpoints)."""
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    out = np.zeros(u.size, dtype=np.int32)
    _lib.check_host(_lib.host().mcb_biconnected_components(int(n), int(u.size), u.ctypes.data, v.ctypes.data,
                                                           out.ctypes.data))
    return out


def vertex_features(row_offsets, cols, weights, alive=None) -> np.ndarray:
    """Per vertex of a CSR multigraph (of the subgraph induced by ``alive`` when given): log(1 + degree),
    log(1 + neighbours' degree sum), core number / max core, mean incident weight / mean weight (float32,
    n x 4; zeros for vertices not alive)."""
    ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
    c = np.ascontiguousarray(cols, dtype=np.uint32)
    wt = np.ascontiguousarray(weights, dtype=np.int32)
    al = None if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
    n = ro.size - 1
    out = np.zeros((n, 4), dtype=np.float32)
    _lib.check_host(_lib.host().mcb_vertex_features(n, ro.ctypes.data, c.ctypes.data, wt.ctypes.data,
                                                    None if al is None else al.ctypes.data, out.ctypes.data))
    return out


def fvs_peel(row_offsets, cols, alive: np.ndarray) -> int:
    """Clears (in place) every alive vertex left with at most one edge to alive vertices, transitively;
    returns the number still alive.  ``alive`` must be a contiguous uint8 array."""
    ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
    c = np.ascontiguousarray(cols, dtype=np.uint32)
    if alive.dtype != np.uint8 or not alive.flags.c_contiguous:
        raise ValueError("alive must be a contiguous uint8 array")
    return _lib.check_host(_lib.host().mcb_fvs_peel(ro.size - 1, ro.ctypes.data, c.ctypes.data, alive.ctypes.data))


class FvsRounds:
    """The learned construction's round state (``mcb_fvs_rounds_*``): the 2-core of the graph minus the
    picks so far, compacted as it shrinks, so a round costs time in what is left."""

    def __init__(self, row_offsets, cols, weights):
        ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
        c = np.ascontiguousarray(cols, dtype=np.uint32)
        wt = np.ascontiguousarray(weights, dtype=np.int32)
        if c.size != wt.size or c.size != int(ro[-1]):
            raise ValueError("cols and weights must hold row_offsets[-1] entries")
        h = ctypes.c_void_p()
        _lib.check_host(_lib.host().mcb_fvs_rounds_create(ro.size - 1, ro.ctypes.data, c.ctypes.data,
                                                          wt.ctypes.data, ctypes.byref(h)))
        self._h = h

    @property
    def left(self) -> int:
        return _lib.check_host(_lib.host().mcb_fvs_rounds_left(self._h))

    def features(self) -> tuple[np.ndarray, np.ndarray]:
        """(ids ascending, their feature rows) of the vertices left."""
        k = self.left
        ids = np.zeros(k, dtype=np.int32)
        X = np.zeros((k, 4), dtype=np.float32)
        _lib.check_host(_lib.host().mcb_fvs_rounds_features(self._h, k, ids.ctypes.data, X.ctypes.data))
        return ids, X

    def take_top(self, scores, k: int) -> np.ndarray:
        """Takes the k best rows of the last ``features`` call by score (ties: lower id), best first."""
        s = np.ascontiguousarray(scores, dtype=np.float32)
        out = np.zeros(min(int(k), s.size), dtype=np.int32)
        _lib.check_host(_lib.host().mcb_fvs_rounds_take_top(self._h, s.ctypes.data, s.size, int(k), out.ctypes.data))
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.host().mcb_fvs_rounds_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fvs_greedy(row_offsets, cols) -> np.ndarray:
    """The peeling max-degree greedy FVS made minimal by reverse-delete; root rank per vertex (-1 = not in F)."""
    ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
    c = np.ascontiguousarray(cols, dtype=np.uint32)
    n = ro.size - 1
    out = np.zeros(n, dtype=np.int32)
    _lib.check_host(_lib.host().mcb_fvs_greedy(n, int(c.size), ro.ctypes.data, c.ctypes.data, out.ctypes.data))
    return out
