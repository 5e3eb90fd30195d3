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
