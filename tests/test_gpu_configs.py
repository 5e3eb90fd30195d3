"""The benchmark configurations at full size on the GPU, each pinned by the oracle's per-phase checks on
sampled phases (tests/fullsize.py) and by the uniqueness of the MCB weight across independent root sets:

* Delaunay 100k (the planar target configuration): sampled phases given the GPU's S_e, the whole selection
  sequence equal for 64 and 7 phases per pass over the trees, and the same weight and dimension with a
  different FVS (the learned construction's root set);
* Chung-Lu 50k with the learned FVS: sampled phases of the learned run, and its weight equal to the plain run's;
* 2000 biconnected components batched in one arena: every component's selections equal its own solver's, and
  small components against the whole-loop oracle;
* a problem with 400 000 coordinates (beyond the per-phase kernel's LDS): the blocked form runs and its
  first blocks pass the sampled checks."""
import os

import numpy as np
import pytest

from fullsize import check_loop, sample_phases
from minimumcyclebasis_amd import _lib
from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.device import Batch, Solver
from minimumcyclebasis_amd.host import Problem
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def scored_order(n, u, v, w) -> np.ndarray:
    """The learned construction's FVS order (verified minimal by the reverse-delete pass), whether or not it
    beats the reference order's size: a root set of its own."""
    from minimumcyclebasis_amd.learned_fvs import default_scorer, graph_of, scored_greedy

    g = Problem.reduced_graph(n, u, v, w)
    try:
        ro, cols, wt = graph_of(g)
        _, order = scored_greedy(ro, cols, wt, default_scorer().logits)
        return order
    finally:
        g.close()


def run_solver(p: Problem, **kw):
    s = Solver(p, **kw)
    s.run()
    s.sync()
    return s


def total_weight(p: Problem, sel: np.ndarray) -> int:
    return int(p.array("cand_weight")[sel].astype(np.int64).sum())


@pytest.fixture(scope="module")
def delaunay():
    gr = G.delaunay(100000, seed=1)
    print("case delaunay100k: host preprocessing", flush=True)
    p = Problem.from_edges(*gr, prune_trees=True, num_threads=threads())
    yield gr, p
    p.close()


def test_delaunay100k_sampled_phases(delaunay):
    gr, p = delaunay
    print(f"case delaunay100k N={p.info.num_non_tree} F={p.info.fvs_size} K={p.info.num_candidates}", flush=True)
    s = run_solver(p)
    sel = s.selected()
    check_loop(p, sel, s.support_vector_rows, sample_phases(p.info.num_non_tree, spread=80), seed=1,
               label="delaunay100k")
    s.close()


def test_delaunay100k_block_sizes_and_root_sets_agree(delaunay):
    gr, p = delaunay
    print("case delaunay100k: 64 and 7 phases per pass", flush=True)
    s = run_solver(p)
    sel = s.selected()
    s.close()
    s7 = run_solver(p, block_phases=7)
    assert np.array_equal(s7.selected(), sel)
    s7.close()
    print("case delaunay100k: learned root set", flush=True)
    pl = Problem.from_edges(*gr, prune_trees=True, num_threads=threads(), fvs_order=scored_order(*gr))
    assert pl.info.num_non_tree == p.info.num_non_tree
    assert not np.array_equal(pl.array("roots"), p.array("roots"))
    sl = run_solver(pl)
    sel_l = sl.selected()
    check_loop(pl, sel_l, sl.support_vector_rows, sample_phases(pl.info.num_non_tree, spread=12), seed=2,
               label="delaunay100k learned")
    assert total_weight(pl, sel_l) == total_weight(p, sel)
    sl.close()
    pl.close()


def test_chunglu50000_learned_and_plain():
    from minimumcyclebasis_amd.learned_fvs import learned_order

    gr = G.chung_lu_lbc(50000, 2.3, 8.0, 10000, seed=1)
    print("case chunglu50000: plain", flush=True)
    p = Problem.from_edges(*gr, prune_trees=True, num_threads=threads())
    s = run_solver(p)
    sel = s.selected()
    s.close()
    ch = learned_order(*gr)
    print(f"case chunglu50000: learned F {ch.fvs_learned} vs reference {ch.fvs_reference}", flush=True)
    assert ch.used and ch.fvs_learned < ch.fvs_reference
    pl = Problem.from_edges(*gr, prune_trees=True, num_threads=threads(), fvs_order=ch.order)
    assert pl.info.fvs_size == ch.fvs_learned
    sl = run_solver(pl)
    sel_l = sl.selected()
    check_loop(pl, sel_l, sl.support_vector_rows, sample_phases(pl.info.num_non_tree, spread=40), seed=3,
               label="chunglu50000 learned")
    assert pl.info.num_non_tree == p.info.num_non_tree
    assert total_weight(pl, sel_l) == total_weight(p, sel)
    sl.close()
    pl.close()
    p.close()


def test_components2000_batched_against_each_solver_and_oracle():
    from minimumcyclebasis_amd.components import build_problems, split_components

    n, u, v, w, _ = G.batched_components(k=2000, nmin=50, nmax=5000, seed=1)
    comps = split_components(n, u, v, w)
    assert len(comps) == 2000
    print("case components2000: host preprocessing", flush=True)
    probs = build_problems(comps, threads=threads(), prune_trees=True)
    loop = [i for i, p in enumerate(probs) if not p.info.single_cycle]
    bt = Batch([probs[i] for i in loop], waves_per_block=4, tree_cap=-1)
    bt.run()
    bt.sync()
    sels = [bt.selected(j) for j in range(len(loop))]
    bt.close()
    print(f"case components2000: {len(loop)} component loops, batched done; one solver each", flush=True)
    for j, i in enumerate(loop):
        s = run_solver(probs[i], waves_per_block=4, tree_cap=-1)
        assert np.array_equal(s.selected(), sels[j]), f"component {i}"
        s.close()
        if j % 500 == 0:
            print(f"case components2000: solver {j}", flush=True)
    # the smallest component loops against the whole-loop oracle
    small = sorted(range(len(loop)), key=lambda j: probs[loop[j]].info.num_non_tree)[:: max(1, len(loop) // 400)][:20]
    for j in small:
        c = comps[loop[j]]
        po = Problem.from_edges(c.n, c.u, c.v, c.w, keep_vertex_trees=True)
        ref = pyoracle.run(po)
        assert np.array_equal(ref["selected"], sels[j]), f"component {loop[j]} against the oracle"
        po.close()
    for p in probs:
        p.close()


def test_beyond_the_per_phase_lds_limit():
    """400 000 coordinates (W = 6250 words: 12 W bytes exceed the per-phase kernel's 64 KB) on two vertices
    joined by 400 001 parallel edges: the per-phase form is refused, the blocked form runs."""
    M = 400001
    rng = np.random.default_rng(4)
    p = Problem.from_edges(2, np.zeros(M, np.int32), np.ones(M, np.int32), rng.integers(1, 1000, size=M),
                           prune_trees=True)
    N = p.info.num_non_tree
    assert N == 400000 and 12 * p.info.words > 65536
    print(f"case parallel edges N={N}", flush=True)
    with pytest.raises(_lib.MCBError):
        Solver(p, block_phases=1)
    s = Solver(p)
    with pytest.raises(_lib.MCBError):
        s.parity(np.zeros(p.info.words, dtype=np.uint64))
    s.run(192)
    s.sync()
    sel = s.selected(192)
    assert np.unique(sel).size == 192
    rows = s.support_vector_rows(np.arange(192, dtype=np.int32))
    for j in range(192):
        bits = np.flatnonzero(np.unpackbits(rows[j].view(np.uint8), bitorder="little"))
        assert bits.max() == j
        assert pyoracle.phase_select(p, rows[j], taken=sel[:j]) == sel[j], f"phase {j}"
    assert not pyoracle.orthogonality(p, sel, np.arange(192), rows, seed=5).any()
    s.close()
    p.close()
