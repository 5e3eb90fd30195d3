"""GPU parity tests of the batched phase loop (mcb_batch_*): many problems packed into one device arena,
each problem's selections and cycle vectors bit-exact against the CPU oracle (mcb/mcb_cycle.cpp:226-305
restated), against the reference's own chosen cycles in the golden traces, and against the problem's own
Solver run alone."""
import numpy as np
import pytest

from helpers import load_trace, problem_of, trace_names
from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.components import split_components
from minimumcyclebasis_amd.device import Batch, Solver, minimum_cycle_basis, minimum_cycle_basis_components
from minimumcyclebasis_amd.host import Problem
from oracle import pyoracle

pytestmark = pytest.mark.gpu

TRACE_CASES = trace_names()


def _check_against_oracle(bt, j, p):
    ref = pyoracle.run(p, want_c=True)
    assert np.array_equal(bt.selected(j), ref["selected"]), f"problem {j}"
    if bt.store_cycles:
        assert np.array_equal(bt.cycle_vectors(j), ref["C"]), f"problem {j}"


@pytest.mark.parametrize("prune", [False, True])
def test_batch_golden_traces(prune):
    """Every golden-trace graph in one batch, in both layouts: per problem, the oracle's selections and
    cycle vectors and the reference's own chosen cycles."""
    cases, probs = [], []
    for name in TRACE_CASES:
        fx = load_trace(name)
        for mode in ("reduced", "plain"):
            p = problem_of(fx, mode, keep_vertex_trees=True, prune_trees=prune)
            if p.info.single_cycle:
                continue
            cases.append((name, mode, fx))
            probs.append(p)
    print(f"case golden batch prune={prune} problems={len(probs)}", flush=True)
    bt = Batch(probs, store_cycles=True)
    bt.run()
    bt.sync()
    assert bt.step == bt.steps == max((p.info.num_non_tree + 63) // 64 for p in probs)
    for j, ((name, mode, fx), p) in enumerate(zip(cases, probs)):
        _check_against_oracle(bt, j, p)
        r, a, b, w = p.candidate_table()
        sel = bt.selected(j)
        chosen = np.stack([r[sel] + 1, a[sel] + 1, b[sel] + 1, w[sel]], axis=1).astype(np.int64).tolist()
        assert chosen == [ph["chosen"] for ph in fx[mode]["phases"]], f"{name} {mode}"
    bt.close()


def _component_problems(k, nmin, nmax, seed, prune=True):
    n, u, v, w, _ = G.batched_components(k=k, nmin=nmin, nmax=nmax, seed=seed)
    probs = []
    for comp in split_components(n, u, v, w):
        p = Problem.from_edges(comp.n, comp.u, comp.v, comp.w, prune_trees=prune, keep_vertex_trees=True)
        if not p.info.single_cycle:
            probs.append(p)
    return probs


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_components_match_solvers_and_oracle(seed):
    """Random 2-connected components of uneven size (several blocks of 64 phases for the largest, one
    partial block for most): each equals its own Solver and the oracle."""
    probs = _component_problems(24, 5, 400, seed)
    print(f"case components seed={seed} problems={len(probs)} N={[p.info.num_non_tree for p in probs]}", flush=True)
    bt = Batch(probs, store_cycles=True)
    bt.run()
    bt.sync()
    for j, p in enumerate(probs):
        s = Solver(p, store_cycles=True)
        s.run()
        s.sync()
        assert np.array_equal(bt.selected(j), s.selected()), f"problem {j}"
        assert np.array_equal(bt.cycle_vectors(j), s.cycle_vectors()), f"problem {j}"
        s.close()
        if p.info.num_non_tree <= 400:
            _check_against_oracle(bt, j, p)
    bt.close()


def test_batch_input_order_and_rerun():
    """Problems given in ascending size (the arena sorts them descending), run in pieces, reset, run again."""
    probs = sorted(_component_problems(10, 20, 300, 7), key=lambda p: p.info.num_non_tree)
    print(f"case rerun problems={len(probs)}", flush=True)
    bt = Batch(probs)
    bt.run(1)
    bt.run(2)
    bt.run()
    bt.sync()
    first = [bt.selected(j).copy() for j in range(len(probs))]
    bt.reset()
    assert bt.step == 0
    assert all((bt.selected(j) == -1).all() for j in range(len(probs)))
    bt.run()
    bt.sync()
    for j, p in enumerate(probs):
        assert np.array_equal(bt.selected(j), first[j])
        _check_against_oracle(bt, j, p)
    bt.close()


def test_batch_single_problem_equals_solver():
    fx = load_trace("grid16x16")
    p = problem_of(fx, "reduced", keep_vertex_trees=True, prune_trees=True)
    print("case single grid16x16", flush=True)
    bt = Batch([p])
    bt.run()
    bt.sync()
    s = Solver(p)
    s.run()
    s.sync()
    assert np.array_equal(bt.selected(0), s.selected())
    pos, cand, upd = bt.work_counters()
    # every candidate is read once per step; a tree at least once per step if it has candidates
    to, rco = p.array("tree_offset"), p.array("root_cand_offset")
    with_cands = int(np.diff(to)[np.diff(rco) > 0].sum())
    assert cand == p.info.num_candidates * bt.steps and pos >= with_cands * bt.steps and upd > 0
    bt.close()
    s.close()


def test_batched_components_driver():
    """minimum_cycle_basis_components(batched=True) against the per-component driver and the whole graph."""
    n, u, v, w, _ = G.batched_components(k=40, nmin=20, nmax=400, seed=11)
    print("case batched driver k=40", flush=True)
    whole = minimum_cycle_basis((n, u, v, w))
    plain = minimum_cycle_basis_components((n, u, v, w))
    batched = minimum_cycle_basis_components((n, u, v, w), batched=True)
    assert batched.total_weight == plain.total_weight == whole.total_weight
    assert batched.num_cycles == plain.num_cycles == len(u) - n + 1
    for (c1, r1), (c2, r2) in zip(plain.components, batched.components):
        assert c1.index == c2.index
        assert np.array_equal(r1.selected, r2.selected)
        assert r2.stats.final_cycles == c2.cycle_space_dim


@pytest.mark.parametrize("cap", [-1, 128])
@pytest.mark.parametrize("waves", [4, 8, 16])
def test_batch_tree_pass_widths(waves, cap):
    probs = _component_problems(12, 20, 600, 3)
    print(f"case batch widths={waves} cap={cap} problems={len(probs)}", flush=True)
    bt = Batch(probs, store_cycles=True, waves_per_block=waves, tree_cap=cap)
    assert bt.launch_shape()["waves_per_block"] == waves
    bt.run()
    bt.sync()
    for j, p in enumerate(probs):
        _check_against_oracle(bt, j, p)
    bt.close()
