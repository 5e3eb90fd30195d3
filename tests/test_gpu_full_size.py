"""Full-size GPU runs: the 64x64 grid against the reference's own stats CSV (Total_Weight,
Final_cycles) and the oracle's phase sequence; the Erdos-Renyi 20000/60000 LBC workload through
size-independent properties and a phase-prefix comparison with the oracle."""
import os

import numpy as np
import pytest

from fullsize import check_loop, sample_phases
from helpers import GOLDEN, edge_digest, graph_of, load_csv
from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.device import Solver, minimum_cycle_basis
from minimumcyclebasis_amd.host import Problem
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prune", [False, True])
def test_grid64_against_reference_csv(prune):
    fx = load_csv("grid64x64")
    n, u, v, w = graph_of(fx)
    print(f"case grid64x64 prune={prune}", flush=True)
    res = minimum_cycle_basis((n, u, v, w), prune_trees=prune)
    assert res.total_weight == fx["mcb_cpu"]["total_weight"] == fx["mcb_cpu_baseline"]["total_weight"]
    assert res.num_cycles == fx["mcb_cpu"]["final_cycles"]
    resp = minimum_cycle_basis((n, u, v, w), reduce_degree_two=False, prune_trees=prune)
    assert resp.total_weight == fx["mcb_cpu_baseline"]["total_weight"]


@pytest.mark.parametrize("block", [1, 0])
def test_grid64_sequence_against_oracle(block):
    fx = load_csv("grid64x64")
    n, u, v, w = graph_of(fx)
    print(f"case grid64x64 sequence block={block}", flush=True)
    p = Problem.from_edges(n, u, v, w, keep_vertex_trees=True, prune_trees=True)
    ref = pyoracle.run(p, want_s=True)
    s = Solver(p, block_phases=block)
    s.run()
    s.sync()
    assert np.array_equal(s.selected(), ref["selected"])
    assert np.array_equal(s.support_vectors(), ref["S"])


@pytest.mark.parametrize("name", ["delaunay2000", "er2000", "chunglu2000"])
def test_medium_graphs_against_reference_csv(name):
    fx = load_csv(name)
    n, u, v, w = graph_of(fx)
    print(f"case {name}", flush=True)
    res = minimum_cycle_basis((n, u, v, w))
    assert res.total_weight == fx["mcb_cpu"]["total_weight"]
    assert res.num_cycles == fx["mcb_cpu"]["final_cycles"]


def test_er20000_properties_and_prefix():
    n, u, v, w = G.erdos_renyi_lbc(20000, 60000, 1000, seed=1)
    print("case er20000 full", flush=True)
    p_full = Problem.from_edges(n, u, v, w, prune_trees=False)
    p_pr = Problem.from_edges(n, u, v, w, prune_trees=True)
    N = p_full.info.num_non_tree
    s1 = Solver(p_full)
    s1.run()
    s1.sync()
    sel1 = s1.selected()
    s1.close()
    s2 = Solver(p_pr)
    s2.run()
    s2.sync()
    sel2 = s2.selected()
    S = s2.support_vectors()
    s2.close()
    print("case er20000 one kernel per phase", flush=True)
    s3 = Solver(p_pr, block_phases=1)
    s3.run()
    s3.sync()
    sel3 = s3.selected()
    assert np.array_equal(s3.support_vectors(), S)
    s3.close()
    assert sel1.size == N and np.unique(sel1).size == N
    assert np.array_equal(sel1, sel2)
    assert np.array_equal(sel1, sel3)
    # sampled phases given the GPU's own S_e (selection, <C_e, S_e> = 1, triangular support) and every
    # later S_j orthogonal to all earlier cycles
    check_loop(p_pr, sel2, lambda ph: S[ph], sample_phases(N, spread=60), seed=7, label="er20000")
    # the first phases against the whole-loop oracle
    p_o = Problem.from_edges(n, u, v, w, keep_vertex_trees=True)
    ref = pyoracle.run(p_o, phases=12)
    assert np.array_equal(sel1[:12], ref["selected"])
    # the reference's own mcb_cpu on the same graph (its stats CSV, when the fixture has been written)
    if os.path.exists(os.path.join(GOLDEN, "csv_er20000.json")):
        fx = load_csv("er20000")
        assert edge_digest(n, u, v, w) == fx["edge_digest"]
        assert int(p_pr.array("cand_weight")[sel2].astype(np.int64).sum()) == fx["mcb_cpu"]["total_weight"]
        assert N == fx["mcb_cpu"]["final_cycles"]
        assert p_pr.info.fvs_size == fx["mcb_cpu"]["fvs_size"]


def test_components_driver_matches_whole_graph():
    from minimumcyclebasis_amd.device import minimum_cycle_basis_components

    n, u, v, w, _ = G.batched_components(k=40, nmin=20, nmax=400, seed=5)
    print("case batched_components k=40", flush=True)
    whole = minimum_cycle_basis((n, u, v, w))
    parts = minimum_cycle_basis_components((n, u, v, w))
    assert parts.total_weight == whole.total_weight
    assert parts.num_cycles == whole.num_cycles == len(u) - n + 1
    assert len(parts.components) == 40
    for comp, res in parts.components:
        assert res.stats.final_cycles == comp.cycle_space_dim
