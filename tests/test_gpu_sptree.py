"""The roots' shortest-path trees searched on the device (mcb_sptree_*) against the host library's own
stopped searches: per root, the accepted candidate edges with their weights and the layout vertices with
their parent edge, level and finalisation index are the same sets; the problem finished from the device
results has every array of the host-built problem.  Cases: the golden graphs of the reference traces,
graphs with many equal distances (the heap's pop order among ties decides the trees), multigraphs from
degree-two contraction (parallel edges relaxed in sequence), hub rows longer than a wave, heaps beyond
the first stage's LDS capacity (later stages and the host fallback), pools that must grow, and the
100 000-point Delaunay triangulation at full size."""
import os

import numpy as np
import pytest

from helpers import graph_of, load_trace, trace_names
from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.device import SpTreeSearch, problem_from_edges
from minimumcyclebasis_amd.host import Problem

pytestmark = pytest.mark.gpu

ARRAYS = ("cand_root", "cand_edge", "cand_weight", "cand_pos_u", "cand_pos_v", "cand_coord", "tree_offset",
          "parent_pos", "pos_coord", "root_cand_offset", "root_cand_rank")


def per_root_equal(p: Problem, res: dict, roots=None) -> None:
    F = p.info.fvs_size
    for r in (range(F) if roots is None else roots):
        if res["status"][r] != 0:
            continue
        e, w, kept = p.root_search(r)
        c0, cn = int(res["cand_start"][r]), int(res["cand_count"][r])
        k0, kn = int(res["kept_start"][r]), int(res["kept_count"][r])
        de, dw = res["cand_edge"][c0:c0 + cn], res["cand_weight"][c0:c0 + cn]
        dk = res["kept"][k0:k0 + kn]
        assert sorted(zip(e.tolist(), w.tolist())) == sorted(zip(de.tolist(), dw.tolist())), f"root {r} candidates"
        assert sorted(map(tuple, kept.tolist())) == sorted(map(tuple, dk.tolist())), f"root {r} layout vertices"


def arrays_equal(a: Problem, b: Problem) -> None:
    for f in ("fvs_size", "num_candidates", "positions", "max_positions", "max_depth", "num_non_tree"):
        assert getattr(a.info, f) == getattr(b.info, f), f
    for name in ARRAYS:
        assert np.array_equal(a.array(name), b.array(name)), name


def check_graph(gr, label, stages=3, per_root=True, expect_stage=None):
    print(f"case {label} stages={stages}", flush=True)
    ref = Problem.from_edges(*gr, prune_trees=True)
    if ref.info.single_cycle:
        ref.close()
        pytest.skip("the whole graph is one cycle: no trees")
    p = Problem.for_external_trees(*gr)
    sp = SpTreeSearch(p)
    sp.run(stages)
    res = sp.results()
    st = sp.stats()
    if per_root:
        per_root_equal(p, res)
    sp.finish(p)
    sp.close()
    arrays_equal(p, ref)
    ref.close()
    p.close()
    return st


@pytest.mark.parametrize("name", trace_names())
def test_golden_graphs(name):
    fx = load_trace(name)
    gr = graph_of(fx)
    if gr[1].size - gr[0] + 1 < 2:
        pytest.skip("no cycle space to search")
    st = check_graph(gr, f"golden {name}")
    assert st["host_roots"] == 0


@pytest.mark.parametrize("gen,label", [
    (lambda: G.grid(30, 30, 2, seed=3), "grid30 weights 1..2"),
    (lambda: G.grid(40, 25, 1, seed=1), "grid40x25 unit weights"),
    (lambda: G.erdos_renyi_lbc(600, 1800, 3, seed=4), "er600 weights 1..3"),
    (lambda: G.delaunay(3000, seed=7), "delaunay3000"),
    (lambda: G.chung_lu_lbc(1500, 2.3, 8.0, 5, seed=2), "chunglu1500 hubs, weights 1..5"),
    (lambda: G.random_biconnected(800, 2.4, 2, seed=5), "biconnected multigraph, weights 1..2"),
    (lambda: G.random_biconnected(400, 2.2, 1, seed=8), "biconnected multigraph, unit weights"),
])
def test_tie_heavy_and_multigraphs(gen, label):
    check_graph(gen(), label)


def test_wheel_hub_rows_longer_than_a_wave():
    """A wheel: the hub's row has every rim vertex (one lane per edge over many chunks)."""
    n = 700
    u = np.concatenate([np.zeros(n - 1, np.int64), np.arange(1, n)])
    v = np.concatenate([np.arange(1, n), np.r_[np.arange(2, n), 1]])
    w = np.random.default_rng(3).integers(1, 4, size=u.size)
    check_graph((n, u, v, w), "wheel700")


def check_staged(gr, label, waves_per_cu, stages):
    print(f"case {label} waves per CU {waves_per_cu} stages={stages}", flush=True)
    ref = Problem.from_edges(*gr, prune_trees=True)
    p = Problem.for_external_trees(*gr)
    sp = SpTreeSearch(p)
    sp.set_stages(waves_per_cu)
    sp.run(stages)
    st = sp.stats()
    print(st, flush=True)
    sp.finish(p)
    sp.close()
    arrays_equal(p, ref)
    ref.close()
    p.close()
    return st


def test_stages_and_host_fallback():
    """Delaunay 60 000 with a first stage of 16 waves per CU (heaps of about 850 entries): about one root in a
    hundred outgrows it.  With one stage those go to the host library (the mixed hand-over), with two they run
    in the second stage; both finish to the host-built arrays."""
    gr = G.delaunay(60000, seed=3)
    st1 = check_staged(gr, "delaunay60000 one stage", [16, 8, 2], 1)
    assert 0 < st1["host_roots"] < st1["stage_roots"][0]
    st2 = check_staged(gr, "delaunay60000 two stages", [16, 8, 2], 2)
    assert st2["stage_roots"][1] > 0 and st2["host_roots"] == 0


def test_expander_goes_to_the_host_after_one_stage():
    """ER 6000 / 18 000: every root's heap outgrows LDS; the first stage finishes under an eighth of the roots,
    so the rest go straight to the host library (no further launches)."""
    gr = G.erdos_renyi_lbc(6000, 18000, 1000, seed=2)
    st = check_staged(gr, "er6000", [8, 2, 1], 3)
    assert st["launches"] == 1 and st["host_roots"] * 8 > 7 * (st["host_roots"] + st["stage_roots"][0])


def test_pools_grow(monkeypatch):
    monkeypatch.setenv("MCB_SPTREE_POOL_MIN", "500")
    check_graph(G.delaunay(2500, seed=4), "delaunay2500 pools from 500 entries")


def test_rerun_and_problem_from_edges():
    gr = G.delaunay(1500, seed=9)
    ref = Problem.from_edges(*gr, prune_trees=True)
    p = Problem.for_external_trees(*gr)
    sp = SpTreeSearch(p)
    sp.run()
    first = sp.results()
    sp.run()
    second = sp.results()
    for k in ("status", "cand_count", "kept_count"):
        assert np.array_equal(first[k], second[k]), k
    per_root_equal(p, second)
    sp.close()
    p.close()
    t = {}
    q = problem_from_edges(*gr, trees="device", timings=t)
    arrays_equal(q, ref)
    assert t["search"]["host_roots"] == 0
    q.close()
    ref.close()


def test_delaunay100k_full_size():
    gr = G.delaunay(100000, seed=1)
    print("case delaunay100k", flush=True)
    th = min(16, len(os.sched_getaffinity(0)))
    ref = Problem.from_edges(*gr, prune_trees=True, num_threads=th)
    t = {}
    p = problem_from_edges(*gr, trees="device", num_threads=th, timings=t)
    print(t, flush=True)
    arrays_equal(p, ref)
    p.close()
    ref.close()
