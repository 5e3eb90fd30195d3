"""Pins the CPU oracle (oracle/mcb_oracle.c) to the reference's own outputs:
per-phase S_e, chosen cycle, C_e and every product <C_e, S_j> on tiny graphs (Test.cpp:239-300),
the chosen-cycle sequence on medium graphs, and Total_Weight / Final_cycles of the stats CSV on
large graphs (include/stats.h:256-338)."""
import numpy as np
import pytest

from helpers import csv_names, dot, graph_of, load_csv, load_trace, problem_of, trace_names, words_of
from minimumcyclebasis_amd.host import Problem
from oracle import pyoracle

MODES = ("reduced", "plain")


def chosen_table(p, sel):
    r, a, b, w = p.candidate_table()
    return np.stack([r[sel] + 1, a[sel] + 1, b[sel] + 1, w[sel]], axis=1).astype(np.int64).tolist()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", trace_names())
def test_chosen_sequence_and_weight(name, mode):
    fx = load_trace(name)
    ref = fx[mode]
    p = problem_of(fx, mode, keep_vertex_trees=True)
    res = pyoracle.run(p, num_threads=2)
    assert chosen_table(p, res["selected"]) == [ph["chosen"] for ph in ref["phases"]]
    assert int(p.array("cand_weight")[res["selected"]].astype(np.int64).sum()) == ref["total_weight"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", trace_names("tiny"))
def test_support_and_cycle_vectors(name, mode):
    fx = load_trace(name)
    ref = fx[mode]
    p = problem_of(fx, mode, keep_vertex_trees=True)
    N, W = p.info.num_non_tree, p.info.words
    res = pyoracle.run(p, num_threads=2, want_s=True, want_c=True)
    # replay the reference's products on its own printed vectors and compare every step
    state = np.zeros((N, W), dtype=np.uint64)
    for j in range(N):
        state[j, j // 64] |= np.uint64(1) << np.uint64(j % 64)
    for e, ph in enumerate(ref["phases"]):
        Se = words_of(ph["S"], W)
        Ce = words_of(ph["C"], W)
        assert np.array_equal(state[e], Se), f"phase {e}: S_e"
        assert np.array_equal(res["S"][e], Se), f"phase {e}: oracle S_e"
        assert np.array_equal(res["C"][e], Ce), f"phase {e}: oracle C_e"
        assert dot(Ce, Se) == 1, "the chosen cycle is odd for S_e"
        for j in range(e + 1, N):
            prod = dot(Ce, state[j])
            assert prod == ph["products"][j - e - 1], f"phase {e}: <C_e, S_{j}>"
            if prod:
                state[j] ^= state[e]


FAST_LARGE = ("grid32x32",)


@pytest.mark.parametrize("name", [pytest.param(n, marks=() if n in FAST_LARGE else pytest.mark.slow)
                                  for n in csv_names() if n != "grid64x64"])
def test_large_total_weight(name):
    fx = load_csv(name)
    n, u, v, w = graph_of(fx)
    for mode, key in (("reduced", "mcb_cpu"), ("plain", "mcb_cpu_baseline")):
        p = Problem.from_edges(n, u, v, w, reduce_degree_two=(mode == "reduced"), keep_vertex_trees=True)
        res = pyoracle.run(p)
        tw = int(p.array("cand_weight")[res["selected"]].astype(np.int64).sum())
        assert tw == fx[key]["total_weight"]
        assert res["selected"].size == fx[key]["final_cycles"]
        assert np.unique(res["selected"]).size == res["selected"].size
        p.close()


@pytest.mark.slow
def test_grid64_total_weight():
    fx = load_csv("grid64x64")
    n, u, v, w = graph_of(fx)
    p = Problem.from_edges(n, u, v, w, keep_vertex_trees=True)
    res = pyoracle.run(p)
    assert int(p.array("cand_weight")[res["selected"]].astype(np.int64).sum()) == fx["mcb_cpu"]["total_weight"]
