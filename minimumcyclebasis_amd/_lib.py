"""ctypes loading of the package's native libraries.

Two shared objects live in ``minimumcyclebasis_amd/lib``:

* ``libmcb_host.so`` - C++ host preprocessing (graph input, degree-two contraction, FVS,
  spanning tree, shortest-path trees, candidate ranking).  Loads on any machine.
* ``libmcb_hip.so``  - the gfx950 HIP kernels and the device C ABI (``include/mcb_hip.h``).

Both are built in-tree by ``minimumcyclebasis_amd.build`` (``__graft_entry__.build()``).  Nothing
here falls back to Python or to the CPU: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
HOST_LIB = os.path.join(LIB_DIR, "libmcb_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libmcb_hip.so")

MCB_OK = 0
ERROR_NAMES = {
    -1: "MCB_E_IO",
    -2: "MCB_E_FORMAT",
    -3: "MCB_E_INVALID",
    -4: "MCB_E_GRAPH",
    -5: "MCB_E_NOMEM",
    -6: "MCB_E_DEVICE",
    -7: "MCB_E_STATE",
    -8: "MCB_E_NOT_FOUND",
}


class MCBError(RuntimeError):
    """A native call returned a negative status."""

    def __init__(self, code: int, message: str):
        self.code = code
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")


_host = None
_hip = None


class BuildOptions(ctypes.Structure):
    _fields_ = [
        ("reduce_degree_two", ctypes.c_int),
        ("prune_trees", ctypes.c_int),
        ("keep_vertex_trees", ctypes.c_int),
        ("num_threads", ctypes.c_int),
        ("fvs_order", ctypes.POINTER(ctypes.c_int32)),
        ("fvs_order_len", ctypes.c_int),
        ("stop_after_fvs", ctypes.c_int),
        ("external_trees", ctypes.c_int),
    ]


class RootTrees(ctypes.Structure):
    """mcb_root_trees (include/mcb_host.h): per-root results of the stopped shortest-path searches."""
    _fields_ = [(name, ctypes.c_void_p) for name in
                ("status", "cand_start", "cand_count", "cand_edge", "cand_weight", "kept_start", "kept_count",
                 "kept")]


class ProblemInfo(ctypes.Structure):
    _fields_ = [
        ("input_nodes", ctypes.c_int32),
        ("input_edges", ctypes.c_int32),
        ("single_cycle", ctypes.c_int32),
        ("nodes_removed", ctypes.c_int32),
        ("source_vertex", ctypes.c_int32),
        ("nodes", ctypes.c_int32),
        ("directed_edges", ctypes.c_int32),
        ("num_non_tree", ctypes.c_int32),
        ("words", ctypes.c_int32),
        ("fvs_size", ctypes.c_int32),
        ("num_candidates", ctypes.c_int64),
        ("positions", ctypes.c_int64),
        ("max_positions", ctypes.c_int32),
        ("max_depth", ctypes.c_int32),
        ("input_total_weight", ctypes.c_int64),
    ]


def _require(path: str, what: str) -> None:
    if not os.path.exists(path):
        raise RuntimeError(
            f"{what} {path} is not built; run __graft_entry__.build() "
            "(or python -m minimumcyclebasis_amd.build) first"
        )


def host() -> ctypes.CDLL:
    """The host preprocessing library (raises if it is not built)."""
    global _host
    if _host is None:
        _require(HOST_LIB, "host library")
        lib = ctypes.CDLL(HOST_LIB)
        P = ctypes.c_void_p
        lib.mcb_build_options_init.argtypes = [ctypes.POINTER(BuildOptions)]
        lib.mcb_build_options_init.restype = None
        lib.mcb_host_last_error.restype = ctypes.c_char_p
        lib.mcb_problem_build_mtx.argtypes = [ctypes.c_char_p, ctypes.POINTER(BuildOptions), ctypes.POINTER(P)]
        lib.mcb_problem_build_edges.argtypes = [ctypes.c_int32, ctypes.c_int32, P, P, P,
                                                ctypes.POINTER(BuildOptions), ctypes.POINTER(P)]
        lib.mcb_problem_get_info.argtypes = [P, ctypes.POINTER(ProblemInfo)]
        lib.mcb_problem_array.argtypes = [P, ctypes.c_int, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_int64)]
        lib.mcb_problem_reference_layout.argtypes = [P, ctypes.c_int32, P, P, P, P]
        lib.mcb_problem_finish_trees.argtypes = [P, ctypes.POINTER(RootTrees), ctypes.c_int]
        lib.mcb_problem_root_search.argtypes = [P, ctypes.c_int32, ctypes.c_int64, P, P, ctypes.c_int64, P,
                                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        lib.mcb_problem_free.argtypes = [P]
        lib.mcb_problem_free.restype = None
        lib.mcb_fvs_reverse_delete.argtypes = [ctypes.c_int32, ctypes.c_int32, P, P, P, P]
        lib.mcb_induced_has_cycle.argtypes = [ctypes.c_int32, P, P, P]
        lib.mcb_biconnected_components.argtypes = [ctypes.c_int32, ctypes.c_int32, P, P, P]
        lib.mcb_vertex_features.argtypes = [ctypes.c_int32, P, P, P, P, P]
        lib.mcb_fvs_peel.argtypes = [ctypes.c_int32, P, P, P]
        lib.mcb_fvs_greedy.argtypes = [ctypes.c_int32, ctypes.c_int32, P, P, P]
        lib.mcb_fvs_rounds_create.argtypes = [ctypes.c_int32, P, P, P, ctypes.POINTER(P)]
        lib.mcb_fvs_rounds_left.argtypes = [P]
        lib.mcb_fvs_rounds_left.restype = ctypes.c_int32
        lib.mcb_fvs_rounds_features.argtypes = [P, ctypes.c_int32, P, P]
        lib.mcb_fvs_rounds_features.restype = ctypes.c_int32
        lib.mcb_fvs_rounds_take_top.argtypes = [P, P, ctypes.c_int32, ctypes.c_int32, P]
        lib.mcb_fvs_rounds_take_top.restype = ctypes.c_int32
        lib.mcb_fvs_rounds_free.argtypes = [P]
        lib.mcb_fvs_rounds_free.restype = None
        _host = lib
    return _host


def hip() -> ctypes.CDLL:
    """The HIP device library (raises if it is not built)."""
    global _hip
    if _hip is None:
        _require(HIP_LIB, "HIP library")
        host()  # the device library links against the host library's symbols
        _hip = ctypes.CDLL(HIP_LIB, mode=ctypes.RTLD_GLOBAL)
        from . import _hipsig

        _hipsig.declare(_hip)
    return _hip


def check_host(rc: int) -> int:
    if rc < 0:
        msg = host().mcb_host_last_error().decode(errors="replace")
        raise MCBError(rc, msg)
    return rc


def check_hip(rc: int) -> int:
    if rc < 0:
        msg = hip().mcb_hip_last_error().decode(errors="replace")
        raise MCBError(rc, msg)
    return rc
