"""Host preprocessing against the reference's own trace output (tests/golden/trace_*.json.gz):
reduced multigraph and its edge order (O1, O2), spanning-tree coordinates (O3, O4), and the ranked
candidate list (O5-O8), in both the degree-two-contracted mode (mcb/Test_deg_two_remove.cpp) and
the plain mode (mcb/Test.cpp, spanning tree at vertex 0)."""
import numpy as np
import pytest

from helpers import load_csv, csv_names, load_trace, problem_of, trace_names, graph_of
from minimumcyclebasis_amd.host import Problem

MODES = ("reduced", "plain")


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", trace_names())
def test_reduced_graph_and_coordinates(name, mode):
    fx = load_trace(name)
    ref = fx[mode]
    p = problem_of(fx, mode)
    info = p.info
    assert info.nodes == ref["nodes"]
    assert info.directed_edges // 2 == ref["edges"]
    rows, cols, w = p.array("rows"), p.array("cols"), p.array("weights")
    mine = [[int(rows[i]) + 1, int(cols[i]) + 1, int(w[i])] for i in range(rows.size) if rows[i] < cols[i]]
    assert mine == ref["reduced_edges"]
    nt = p.array("nt_edges")
    assert info.num_non_tree == ref["N"] == ref["num_non_tree"]
    assert [[int(rows[e]) + 1, int(cols[e]) + 1] for e in nt] == ref["non_tree_edges"]
    assert [[i, int(rows[e]) + 1, int(cols[e]) + 1] for i, e in enumerate(nt)] == ref["nt_map"]
    # both directions of a coordinate edge carry its index; tree edges carry -1
    ntm = p.array("nt_map")
    rev = p.array("reverse")
    assert np.array_equal(ntm, ntm[rev])
    assert np.count_nonzero(ntm >= 0) == 2 * info.num_non_tree


@pytest.mark.parametrize("prune", [False, True])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", trace_names())
def test_candidate_ranking(name, mode, prune):
    """The ranked candidates, from the whole shortest-path trees and from the searches that stop once
    the rest of the graph is blocked (the pruned layout)."""
    fx = load_trace(name)
    ref = fx[mode]
    p = problem_of(fx, mode, prune_trees=prune)
    # mcb/Test.cpp prints no count line; the listed candidates are the count there
    assert p.info.num_candidates == ref.get("num_candidates", len(ref["candidates"])) == len(ref["candidates"])
    r, a, b, w = p.candidate_table()
    mine = np.stack([r + 1, a + 1, b + 1, w], axis=1).astype(np.int64).tolist()
    assert mine == ref["candidates"]


@pytest.mark.parametrize("name", csv_names())
def test_large_counts(name):
    fx = load_csv(name)
    n, u, v, w = graph_of(fx)
    for mode, key in (("reduced", "mcb_cpu"), ("plain", "mcb_cpu_baseline")):
        ref = fx[key]
        p = Problem.from_edges(n, u, v, w, reduce_degree_two=(mode == "reduced"))
        assert p.info.fvs_size == ref["fvs_size"]
        assert p.info.num_candidates == ref["initial_cycles"]
        assert p.info.num_non_tree == ref["final_cycles"]
        assert p.info.nodes_removed == ref["nodes_removed"]
        if mode == "reduced":
            assert p.info.directed_edges == ref["new_edges"]
        p.close()


@pytest.mark.parametrize("name", ["grid16x16", "delaunay150", "chunglu300", "er150"])
def test_pruned_layout_keeps_candidates(name):
    """Pruning keeps exactly the root-to-endpoint closure; ranks and weights do not change."""
    fx = load_trace(name)
    full = problem_of(fx, "reduced", keep_vertex_trees=True)
    pr = problem_of(fx, "reduced", prune_trees=True, keep_vertex_trees=True)
    for key in ("cand_root", "cand_edge", "cand_weight", "cand_coord"):
        assert np.array_equal(full.array(key), pr.array(key))
    assert pr.info.positions <= full.info.positions
    # every kept position's parent is kept and precedes it; endpoints map to the same vertex
    to, par, pv = pr.array("tree_offset"), pr.array("parent_pos"), pr.array("pos_vertex")
    rows, cols = pr.array("rows"), pr.array("cols")
    for k in range(pr.info.num_candidates):
        rt = pr.array("cand_root")[k]
        base = to[rt]
        e = pr.array("cand_edge")[k]
        assert pv[base + pr.array("cand_pos_u")[k]] == rows[e]
        assert pv[base + pr.array("cand_pos_v")[k]] == cols[e]
    for rt in range(pr.info.fvs_size):
        seg = par[to[rt]:to[rt + 1]]
        assert seg[0] == 0xFFFFFFFF
        assert np.all(seg[1:] < np.arange(1, seg.size))


def _stop_cases():
    from minimumcyclebasis_amd import generators as G
    return {
        "delaunay3000": lambda: G.delaunay(3000, seed=5),          # small integer weights: many distance ties
        "grid40x30": lambda: G.grid(40, 30, 5, seed=6),            # weights 1..5: ties everywhere
        "er3000": lambda: G.erdos_renyi_lbc(3000, 9000, 1000, seed=7),
        "chunglu4000": lambda: G.chung_lu_lbc(4000, 2.3, 8.0, 10000, seed=8),
        "grid_unit": lambda: G.grid(25, 25, 1, seed=9),             # all weights equal
    }


@pytest.mark.parametrize("case", ["delaunay3000", "grid40x30", "er3000", "chunglu4000", "grid_unit"])
@pytest.mark.parametrize("reduce", [True, False])
def test_stopped_searches_give_the_same_problem(case, reduce):
    """The pruned layout is built from searches that stop once every unfinished vertex is blocked;
    with per-vertex parent rows requested the searches run to the end.  Every array of the two
    problems agrees: candidates, their ranks and weights, and the position layouts."""
    gr = _stop_cases()[case]()
    full = Problem.from_edges(*gr, reduce_degree_two=reduce, prune_trees=True, keep_vertex_trees=True, num_threads=4)
    stop = Problem.from_edges(*gr, reduce_degree_two=reduce, prune_trees=True, num_threads=4)
    assert stop.info.num_candidates == full.info.num_candidates > 0
    for key in ("cand_root", "cand_edge", "cand_weight", "cand_pos_u", "cand_pos_v", "cand_coord", "tree_offset",
                "parent_pos", "pos_coord", "root_cand_offset", "root_cand_rank"):
        assert np.array_equal(stop.array(key), full.array(key)), key
    full.close()
    stop.close()
