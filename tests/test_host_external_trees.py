"""The problem assembled from per-root search results (the form the device search hands over) against the
problem whose trees the host library searches itself: every array identical.  The results come from the
host's own stopped search (mcb_problem_root_search), handed over in shuffled order with some roots
withheld (status != 0: searched by the host inside mcb_problem_finish_trees), so the order-independence of
the hand-over and the mixed path are covered without a GPU."""
import numpy as np
import pytest

from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.host import Problem

ARRAYS = ("cand_root", "cand_edge", "cand_weight", "cand_pos_u", "cand_pos_v", "cand_coord", "tree_offset",
          "parent_pos", "pos_coord", "root_cand_offset", "root_cand_rank")

GRAPHS = {
    "grid_ties": lambda: G.grid(14, 14, 3, seed=4),
    "er_ties": lambda: G.erdos_renyi_lbc(300, 900, 4, seed=2),
    "delaunay": lambda: G.delaunay(600, seed=5),
    "chunglu": lambda: G.chung_lu_lbc(500, 2.3, 6.0, 20, seed=1),
    "biconn_multi": lambda: G.random_biconnected(200, 2.6, 3, seed=9),
}


def hand_over(p: Problem, withheld=(), seed=0):
    """Per-root results of every root, each root's lists shuffled and the roots stored in shuffled order."""
    rng = np.random.default_rng(seed)
    F = p.info.fvs_size
    res = [p.root_search(r) for r in range(F)]
    status = np.zeros(F, dtype=np.int32)
    status[list(withheld)] = 1
    cs, cc, ks, kc = (np.zeros(F, dtype=np.int64), np.zeros(F, dtype=np.int32), np.zeros(F, dtype=np.int64),
                      np.zeros(F, dtype=np.int32))
    ce, cw, kept = [], [], []
    nc = nk = 0
    for r in rng.permutation(F):
        e, w, k = res[r]
        pc = rng.permutation(e.size)
        pk = rng.permutation(k.shape[0])
        cs[r], cc[r], ks[r], kc[r] = nc, e.size, nk, k.shape[0]
        ce.append(e[pc])
        cw.append(w[pc])
        kept.append(k[pk])
        nc += e.size
        nk += k.shape[0]
    cat = lambda xs, dt, shape: np.concatenate(xs).astype(dt) if xs else np.zeros(shape, dt)  # noqa: E731
    return status, cs, cc, cat(ce, np.uint32, 0), cat(cw, np.int32, 0), ks, kc, cat(kept, np.uint32, (0, 4))


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_assembled_from_results_equals_host_trees(name):
    gr = GRAPHS[name]()
    ref = Problem.from_edges(*gr, prune_trees=True)
    ext = Problem.for_external_trees(*gr)
    assert ext.info.num_candidates == 0 and ext.info.fvs_size == ref.info.fvs_size
    F = ext.info.fvs_size
    withheld = list(range(0, F, 7))
    ext.finish_trees(*hand_over(ext, withheld, seed=F))
    for f in ("num_candidates", "positions", "max_positions", "max_depth", "num_non_tree"):
        assert getattr(ext.info, f) == getattr(ref.info, f), f
    for a in ARRAYS:
        assert np.array_equal(ext.array(a), ref.array(a)), a
    ref.close()
    ext.close()


def test_root_search_is_the_pruned_layout():
    """Each root's kept vertices are exactly its layout's positions past the root, their count the layout
    size, and the accepted edges its candidates."""
    gr = G.delaunay(400, seed=2)
    p = Problem.from_edges(*gr, prune_trees=True)
    to = p.array("tree_offset")
    rco = p.array("root_cand_offset")
    for r in range(p.info.fvs_size):
        e, w, kept = p.root_search(r)
        assert kept.shape[0] + 1 == to[r + 1] - to[r]
        assert e.size == rco[r + 1] - rco[r]
        ranks = p.array("root_cand_rank")[rco[r]:rco[r + 1]]
        assert sorted(p.array("cand_edge")[ranks].tolist()) == sorted(e.tolist())
        assert sorted(p.array("cand_weight")[ranks].tolist()) == sorted(w.tolist())
    p.close()


def test_finish_requires_external_build():
    gr = G.grid(6, 6, 10, seed=1)
    p = Problem.from_edges(*gr, prune_trees=True)
    F = p.info.fvs_size
    z = np.zeros(F, dtype=np.int64)
    with pytest.raises(RuntimeError):
        p.finish_trees(np.ones(F, np.int32), z, z.astype(np.int32), [], [], z, z.astype(np.int32),
                       np.zeros((0, 4), np.uint32))
    p.close()
