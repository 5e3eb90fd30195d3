"""ctypes binding of ``oracle/liboracle.so`` (a C restatement of the reference's CPU phase loop,
mcb/mcb_cycle.cpp:226-305).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (make -C oracle oracle)")
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        L.oracle_parity.argtypes = [i32, i32, P, P, P, P, P, i32, P, P, i32]
        L.oracle_run.argtypes = [i32, i32, P, P, P, P, P, i32, i64, P, P, i32, i32, P, P, P]
        L.oracle_position_parity.argtypes = [i32, P, P, P, i32, P, P, i32]
        L.oracle_phase_select.argtypes = [i32, P, P, P, i64, P, P, P, P, i32, P, P, i32]
        L.oracle_phase_select.restype = i64
        L.oracle_cycle_vector.argtypes = [P, P, P, i64, P, P, P, P, i64, i32, P]
        L.oracle_orthogonality.argtypes = [P, P, P, i64, P, P, P, P, i32, P, P, i32, P, i32, P, P]
        L.oracle_phase_timed.argtypes = [i32, i32, P, P, P, P, P, i32, i64, P, P, P, i32, i32, P, P, P, P, i64, i32,
                                         P]
        L.oracle_phase_timed.restype = i64
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def parity(problem, S: np.ndarray, num_threads: int = 0) -> np.ndarray:
    """Per-vertex tree-path parity P[r, v] of every root for support vector S (uint64 words)."""
    info = problem.info
    n, F = info.nodes, info.fvs_size
    vp = problem.array("vertex_parent")
    if vp.size != n * F:
        raise ValueError("problem was built without keep_vertex_trees")
    S = _c(S, np.uint64)
    out = np.zeros((F, n), dtype=np.uint8)
    rc = lib().oracle_parity(n, F, problem.array("roots").ctypes.data, problem.array("rows").ctypes.data,
                             problem.array("cols").ctypes.data, vp.ctypes.data, problem.array("nt_map").ctypes.data,
                             int(S.size), S.ctypes.data, out.ctypes.data, int(num_threads))
    if rc:
        raise RuntimeError(f"oracle_parity failed: {rc}")
    return out


def run(problem, phases: int | None = None, num_threads: int = 0, want_s: bool = False, want_c: bool = False):
    """Runs the phase loop; returns dict(selected, S (N x W) or None, C (phases x W) or None)."""
    info = problem.info
    n, F, N, K = info.nodes, info.fvs_size, info.num_non_tree, info.num_candidates
    W = info.words
    phases = N if phases is None else int(phases)
    vp = problem.array("vertex_parent")
    if vp.size != n * F:
        raise ValueError("problem was built without keep_vertex_trees")
    sel = np.full(phases, -1, dtype=np.int32)
    S = np.zeros((N, W), dtype=np.uint64) if want_s else None
    C = np.zeros((phases, W), dtype=np.uint64) if want_c else None
    rc = lib().oracle_run(n, F, problem.array("roots").ctypes.data, problem.array("rows").ctypes.data,
                          problem.array("cols").ctypes.data, vp.ctypes.data, problem.array("nt_map").ctypes.data,
                          N, K, problem.array("cand_root").ctypes.data, problem.array("cand_edge").ctypes.data,
                          phases, int(num_threads), sel.ctypes.data,
                          None if S is None else S.ctypes.data, None if C is None else C.ctypes.data)
    if rc:
        raise RuntimeError(f"oracle_run failed: {rc}")
    return {"selected": sel, "S": S, "C": C}


def phase_timed(problem, e: int, S_e: np.ndarray, taken, P: np.ndarray, later: np.ndarray, root_sample,
                num_threads: int = 0):
    """One phase of the reference's CPU loop from S_e, timed in parts: returns (chosen rank, seconds of
    the parity BFS of the ``root_sample`` roots, of the list scan over ``P`` (F x n uint8, per-vertex
    parity of every root, filled by the caller), and of the cycle vector plus the products with (and
    XORs into) the rows of ``later`` (modified in place))."""
    info = problem.info
    n, F, N, K, W = info.nodes, info.fvs_size, info.num_non_tree, info.num_candidates, info.words
    vp = problem.array("vertex_parent")
    if vp.size != n * F:
        raise ValueError("problem was built without keep_vertex_trees")
    if P.dtype != np.uint8 or P.size != n * F or not P.flags.c_contiguous:
        raise ValueError("P must be a contiguous uint8 array of F x n")
    if later.dtype != np.uint64 or not later.flags.c_contiguous or (later.size and later.shape[-1] != W):
        raise ValueError("later must be contiguous uint64 rows of W words")
    S_e = _c(S_e, np.uint64)
    rs = _c(root_sample, np.int32)
    tk = np.zeros(K, dtype=np.uint8)
    if taken is not None and len(taken):
        tk[np.asarray(taken, dtype=np.int64)] = 1
    t = np.zeros(3, dtype=np.float64)
    k = lib().oracle_phase_timed(n, F, problem.array("roots").ctypes.data, problem.array("rows").ctypes.data,
                                 problem.array("cols").ctypes.data, vp.ctypes.data,
                                 problem.array("nt_map").ctypes.data, N, K, problem.array("cand_root").ctypes.data,
                                 problem.array("cand_edge").ctypes.data, rs.ctypes.data, int(rs.size), W,
                                 S_e.ctypes.data, tk.ctypes.data, P.ctypes.data,
                                 later.ctypes.data if later.size else None, int(later.size // W if W else 0),
                                 int(num_threads), t.ctypes.data)
    if k < 0:
        raise RuntimeError(f"oracle_phase_timed failed: {k}")
    return int(k), float(t[0]), float(t[1]), float(t[2])


# ---------------------------------------------------------------------------------------------
# Per-phase checks on the position layout (tree_offset / parent_pos / pos_coord and the rank-ordered
# candidate arrays of a Problem), for graphs too large for the per-vertex rows above.
# ---------------------------------------------------------------------------------------------

def _layout(problem):
    return [problem.array(n) for n in ("tree_offset", "parent_pos", "pos_coord")]


def _cands(problem):
    return [problem.array(n) for n in ("cand_root", "cand_pos_u", "cand_pos_v", "cand_coord")]


def position_parity(problem, S: np.ndarray, num_threads: int = 0) -> np.ndarray:
    """Parity of every position of every root's layout for S (uint8, one per position)."""
    to, pp, pc = _layout(problem)
    S = _c(S, np.uint64)
    out = np.zeros(problem.info.positions, dtype=np.uint8)
    rc = lib().oracle_position_parity(problem.info.fvs_size, to.ctypes.data, pp.ctypes.data, pc.ctypes.data,
                                      int(S.size), S.ctypes.data, out.ctypes.data, int(num_threads))
    if rc:
        raise RuntimeError(f"oracle_position_parity failed: {rc}")
    return out


def phase_select(problem, S: np.ndarray, taken=None, num_threads: int = 0) -> int:
    """Rank of the first candidate (in rank order, skipping ``taken`` ranks) whose cycle is odd for S;
    -1 when none is."""
    to, pp, pc = _layout(problem)
    cr, cu, cv, cc = _cands(problem)
    S = _c(S, np.uint64)
    K = problem.info.num_candidates
    tk = None
    if taken is not None:
        tk = np.zeros(K, dtype=np.uint8)
        tk[np.asarray(taken, dtype=np.int64)] = 1
    k = lib().oracle_phase_select(problem.info.fvs_size, to.ctypes.data, pp.ctypes.data, pc.ctypes.data, K,
                                  cr.ctypes.data, cu.ctypes.data, cv.ctypes.data, cc.ctypes.data, int(S.size),
                                  S.ctypes.data, None if tk is None else tk.ctypes.data, int(num_threads))
    if k < -1:
        raise RuntimeError(f"oracle_phase_select failed: {k}")
    return int(k)


def cycle_vector(problem, rank: int) -> np.ndarray:
    """Cycle vector (W words) of the candidate of this rank."""
    to, pp, pc = _layout(problem)
    cr, cu, cv, cc = _cands(problem)
    W = problem.info.words
    out = np.zeros(W, dtype=np.uint64)
    rc = lib().oracle_cycle_vector(to.ctypes.data, pp.ctypes.data, pc.ctypes.data, problem.info.num_candidates,
                                   cr.ctypes.data, cu.ctypes.data, cv.ctypes.data, cc.ctypes.data, int(rank), W,
                                   out.ctypes.data)
    if rc:
        raise RuntimeError(f"oracle_cycle_vector failed: {rc}")
    return out


def orthogonality(problem, selected: np.ndarray, samples, S_samples: np.ndarray, seed: int = 0) -> np.ndarray:
    """For each sample phase j (ascending), 64 random GF(2) combinations of <C_i, S_j> over every i < j
    (C_i the cycle vector of rank selected[i]); all zero when S_j is orthogonal to every earlier cycle."""
    to, pp, pc = _layout(problem)
    cr, cu, cv, cc = _cands(problem)
    N, W = problem.info.num_non_tree, problem.info.words
    sel = _c(selected, np.int32)
    if sel.size < N:
        sel = np.concatenate([sel, np.zeros(N - sel.size, dtype=np.int32)])
    samp = _c(samples, np.int32)
    if samp.size and int(samp.max()) > 0 and int(sel[: int(samp.max())].min(initial=0)) < 0:
        raise ValueError("a phase before a sample has no selection")
    Ss = _c(S_samples, np.uint64).reshape(samp.size, W)
    r = np.random.default_rng(seed).integers(0, 2**64, size=N, dtype=np.uint64)
    out = np.zeros(samp.size, dtype=np.uint64)
    rc = lib().oracle_orthogonality(to.ctypes.data, pp.ctypes.data, pc.ctypes.data, problem.info.num_candidates,
                                    cr.ctypes.data, cu.ctypes.data, cv.ctypes.data, cc.ctypes.data, N,
                                    sel.ctypes.data, r.ctypes.data, int(samp.size), samp.ctypes.data, W,
                                    Ss.ctypes.data, out.ctypes.data)
    if rc:
        raise RuntimeError(f"oracle_orthogonality failed: {rc}")
    return out
