"""Host side of the boundary, on the CPU: every function include/*.h declares is exported by its
library; the union-find FVS equals a literal restatement of the reference's MGA + reverse-delete
(include/FVS.cpp:166-214); input errors come back as status codes, not crashes; the generators are
deterministic."""
import ctypes
import os
import re

import numpy as np
import pytest

from minimumcyclebasis_amd import _lib
from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.host import Problem, fvs_reverse_delete, induced_has_cycle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header: str) -> list[str]:
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"#[^\n]*", "", text)
    names = []
    for decl in text.split(";"):
        decl = " ".join(decl.split())
        decl = re.split(r"[{}]", decl)[-1].strip()
        m = re.match(r"^(?:const\s+)?[A-Za-z_][\w]*[\s\*]+([A-Za-z_]\w*)\s*\(", decl)
        if m:
            names.append(m.group(1))
    return names


@pytest.mark.parametrize("header,lib", [("mcb_host.h", _lib.HOST_LIB), ("mcb_hip.h", _lib.HIP_LIB),
                                        ("mcb_mlp.h", _lib.HIP_LIB)])
def test_header_symbols_exported(header, lib):
    names = declared_functions(header)
    assert len(names) >= {"mcb_host.h": 10, "mcb_hip.h": 30, "mcb_mlp.h": 9}[header], names
    so = ctypes.CDLL(lib)  # loading needs no GPU
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{os.path.basename(lib)} lacks {missing}"


def test_reference_entry_points_declared():
    names = set(declared_functions("mcb_hip.h"))
    for ref in ("init_cuda", "configure_grid", "allocate_pinned_memory", "free_pinned_memory",
                "calculate_chunk_size"):
        assert ref in names


def test_device_calls_fail_loudly_without_gpu():
    from minimumcyclebasis_amd import device

    if device.device_count() > 0:
        pytest.skip("a GPU is visible")
    n, u, v, w = G.grid(4, 4, 5, seed=1)
    p = Problem.from_edges(n, u, v, w, prune_trees=True)
    with pytest.raises(_lib.MCBError):
        device.Solver(p)


# --- FVS: literal restatement of FVS::MGA (include/FVS.cpp:166-214) ---------------------------
# MGA picks, while any vertex is left, the first vertex of minimum W/deg with W = 1 (pruning only
# retires the picked vertex, so degrees never change): that is degree descending, id ascending.
# The list is reversed and each vertex is dropped from the set if the rest stays acyclic
# (test_fvs; a pair of parallel edges is a cycle).  Acyclicity here is edges == vertices - components.

def _forest(n, rows, cols, keep_out):
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    verts = [x for x in range(n) if not keep_out[x]]
    edges = 0
    comps = len(verts)
    for a, b in zip(rows, cols):
        if a < b and not keep_out[a] and not keep_out[b]:
            edges += 1
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[ra] = rb
                comps -= 1
        elif a == b and not keep_out[a]:
            return False
    return edges == len(verts) - comps


def _mga_reverse_delete(n, row_offsets, rows, cols):
    deg = np.diff(row_offsets.astype(np.int64))
    picked = sorted(range(n), key=lambda x: (-int(deg[x]), x))
    in_fvs = [True] * n
    for x in reversed(picked):
        in_fvs[x] = False
        if not _forest(n, rows, cols, in_fvs):
            in_fvs[x] = True
    return np.array(in_fvs)


@pytest.mark.parametrize("gen", [
    lambda: G.grid(6, 7, 9, seed=1),
    lambda: G.complete(7, 9, seed=2),
    lambda: G.erdos_renyi_lbc(60, 150, 20, seed=3),
    lambda: G.delaunay(70, seed=5),
    lambda: G.chung_lu_lbc(120, 2.3, 6.0, 40, seed=4),
])
@pytest.mark.parametrize("mode", ("reduced", "plain"))
def test_fvs_matches_literal_reverse_delete(gen, mode):
    n, u, v, w = gen()
    p = Problem.from_edges(n, u, v, w, reduce_degree_two=(mode == "reduced"))
    ro, rows, cols = p.array("row_offsets"), p.array("rows"), p.array("cols")
    nn = p.info.nodes
    want = _mga_reverse_delete(nn, ro, rows.tolist(), cols.tolist())
    got = p.array("fvs_rank") >= 0
    assert np.array_equal(got, want)
    # ranks are the vertex ids of the set in ascending order (FVS.cpp:216-227)
    assert np.array_equal(p.array("fvs_rank")[got], np.arange(got.sum()))
    assert np.array_equal(fvs_reverse_delete(ro, cols) >= 0, want)
    assert not induced_has_cycle(ro, cols, want.astype(np.uint8))
    # dropping any kept vertex from the set leaves a cycle (the set is minimal)
    for x in np.flatnonzero(want)[:10]:
        s = want.copy()
        s[x] = False
        assert induced_has_cycle(ro, cols, s.astype(np.uint8))


def test_fvs_custom_order_is_verified_and_minimal():
    n, u, v, w = G.erdos_renyi_lbc(80, 200, 20, seed=8)
    p = Problem.from_edges(n, u, v, w)
    ro, cols = p.array("row_offsets"), p.array("cols")
    order = np.random.default_rng(1).permutation(p.info.nodes).astype(np.int32)
    rank = fvs_reverse_delete(ro, cols, order)
    in_f = rank >= 0
    assert not induced_has_cycle(ro, cols, in_f.astype(np.uint8))
    for x in np.flatnonzero(in_f):
        s = in_f.copy()
        s[x] = False
        assert induced_has_cycle(ro, cols, s.astype(np.uint8))


# --- error behaviour -----------------------------------------------------------------------------

def test_missing_file(tmp_path):
    with pytest.raises(_lib.MCBError) as ei:
        Problem.from_mtx(str(tmp_path / "nope.mtx"))
    assert ei.value.code == -1


def test_bad_banner(tmp_path):
    f = tmp_path / "bad.mtx"
    f.write_text("this is not matrix market\n3 3 3\n1 2 1\n2 3 1\n3 1 1\n")
    with pytest.raises(_lib.MCBError):
        Problem.from_mtx(str(f))


def test_edge_endpoint_out_of_range():
    with pytest.raises(_lib.MCBError) as ei:
        Problem.from_edges(3, [0, 1, 2], [1, 2, 5], [1, 1, 1])
    assert ei.value.code == -3


def test_mtx_round_trip(tmp_path):
    n, u, v, w = G.grid(5, 6, 9, seed=4)
    f = tmp_path / "g.mtx"
    G.write_mtx(str(f), n, u, v, w)
    a = Problem.from_mtx(str(f))
    b = Problem.from_edges(n, u, v, w)
    for name in ("rows", "cols", "weights", "nt_edges", "fvs_rank", "cand_weight", "cand_edge", "cand_root"):
        assert np.array_equal(a.array(name), b.array(name)), name


def test_single_cycle_graph():
    n = 7
    u = np.arange(n)
    v = (u + 1) % n
    w = np.arange(1, n + 1)
    p = Problem.from_edges(n, u, v, w)
    assert p.info.single_cycle
    assert p.info.input_total_weight == int(w.sum())


@pytest.mark.parametrize("gen,args", [
    ("grid", (7, 5, 30)), ("erdos_renyi_lbc", (200, 600, 100)), ("delaunay", (300,)),
    ("chung_lu_lbc", (400, 2.3, 8.0, 1000)), ("random_biconnected", (120, 5.0, 100)),
])
def test_generators_deterministic(gen, args):
    f = getattr(G, gen)
    a = f(*args, seed=3)
    b = f(*args, seed=3)
    c = f(*args, seed=4)
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)
    assert a[0] == b[0]
    assert not all(np.array_equal(x, y) for x, y in zip(a[1:], c[1:])) or a[0] != c[0]


# --- stats CSV (include/stats.h:256-338) --------------------------------------------------------

@pytest.mark.parametrize("name", ["grid32x32", "er2000", "delaunay2000"])
def test_stats_row_matches_reference_csv(name, tmp_path):
    from helpers import graph_of, load_csv
    from minimumcyclebasis_amd.stats import CPU_COLUMNS, GPU_COLUMNS, append_stats_csv, read_stats_csv, stats_of

    fx = load_csv(name)
    n, u, v, w = graph_of(fx)
    ref = fx["mcb_cpu"]
    p = Problem.from_edges(n, u, v, w)
    st = stats_of(p, ref["total_weight"], construction_trees_s=1.5, inspection_s=2.25)
    for gpu in (False, True):
        path = tmp_path / f"stats_{gpu}.csv"
        append_stats_csv(str(path), st, gpu=gpu)
        append_stats_csv(str(path), st, gpu=gpu)  # a second run appends a row, not a header
        rows = read_stats_csv(str(path))
        assert len(rows) == 2
        row = rows[0]
        assert list(row) == list(GPU_COLUMNS if gpu else CPU_COLUMNS)
        got = [row["Total_Nodes"], row["Total_Edges"], row["New_Edges"],
               row["Nodes_removed" if gpu else "Node_removed"], row["Fvs_size"], row["Initial_Cycles"],
               row["Final_cycles"], row["Total_Weight"]]
        want = [ref[k] for k in ("total_nodes", "total_edges", "new_edges", "nodes_removed", "fvs_size",
                                 "initial_cycles", "final_cycles", "total_weight")]
        assert got == want
        assert row["Total_time(s)"] == pytest.approx(3.75)


def test_stats_int32_weight_wraps_like_the_reference():
    from minimumcyclebasis_amd.stats import RunStats

    st = RunStats(1, 2, 2, 0, 1, 1, 1, 2**31 + 5)
    assert int(st.row(False).split(",")[7]) == -(2**31) + 5
