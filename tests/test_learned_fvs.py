"""Learned FVS ordering, CPU side: the MLP oracle pinned by finite differences, the vertex features
against a direct restatement, the greedy teacher (acyclic complement, minimal, smaller than the reference
order's), and the ordering / fallback logic.  The device kernels are checked against the oracle in
test_gpu_learned_fvs.py."""
import numpy as np
import pytest

from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.host import (Problem, fvs_greedy, fvs_peel, fvs_reverse_delete, induced_has_cycle,
                                        vertex_features)
from minimumcyclebasis_amd.learned_fvs import (LearnedChoice, choose_order, degree_score, reference_order,
                                               scored_greedy, training_rows)
from oracle import mlp


def _params(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return (rng.normal(0, scale, mlp.N_PARAMS)).astype(np.float64)


def test_oracle_gradient_matches_finite_differences():
    """The exact (float64) oracle's backward pass against central differences of its loss."""
    rng = np.random.default_rng(1)
    X = rng.normal(0, 1, (24, 4))
    y = (rng.random(24) < 0.4).astype(np.float64)
    p = _params(2)
    _, g = mlp.loss_and_grad(p, X, y, emulate_bf16=False)
    eps = 1e-6
    idx = np.concatenate([rng.choice(256, 12, replace=False), 256 + rng.choice(64, 6, replace=False),
                          320 + rng.choice(4096, 24, replace=False), 4416 + rng.choice(64, 6, replace=False),
                          4480 + rng.choice(64, 6, replace=False), [4544]])
    for i in idx:
        d = np.zeros_like(p)
        d[i] = eps
        num = (mlp.loss_and_grad(p + d, X, y, emulate_bf16=False)[0] -
               mlp.loss_and_grad(p - d, X, y, emulate_bf16=False)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_oracle_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-8], dtype=np.float32)
    b = mlp.bf16(x)
    assert b[0] == 1.0 and b[1] == 1.0  # halfway rounds to even
    assert b[2] == np.float32(1.0078125)
    assert np.all(np.abs(b - x) <= np.abs(x) * 2.0 ** -8)
    # emulated and exact functions agree to BF16 precision
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (50, 4))
    p = _params(4, 0.2)
    z0 = mlp.forward(p, X, emulate_bf16=False)
    z1 = mlp.forward(p, X, emulate_bf16=True)
    assert np.max(np.abs(z0 - z1)) < 0.05 * (1 + np.max(np.abs(z0)))


def _features_direct(ro, cols, w):
    n = ro.size - 1
    deg = np.diff(ro).astype(np.int64)
    two = np.array([deg[cols[ro[v]:ro[v + 1]]].sum() for v in range(n)])
    meanw = np.array([w[ro[v]:ro[v + 1]].mean() if deg[v] else 0.0 for v in range(n)])
    # core numbers by repeated peeling of the least remaining degree
    d = deg.copy()
    alive = np.ones(n, bool)
    core = np.zeros(n, np.int64)
    k = 0
    while alive.any():
        v = np.flatnonzero(alive)[np.argmin(d[alive])]
        k = max(k, d[v])
        core[v] = k
        alive[v] = False
        for u in cols[ro[v]:ro[v + 1]]:
            if alive[u]:
                d[u] -= 1
    return np.stack([np.log1p(deg), np.log1p(two), core / max(1, core.max()), meanw / w.mean()], axis=1)


@pytest.mark.parametrize("case", ["grid", "chunglu", "delaunay", "er"])
def test_vertex_features_direct(case):
    gr = {"grid": lambda: G.grid(8, 9, 50, seed=1), "chunglu": lambda: G.chung_lu_lbc(400, 2.3, 8.0, 10000, seed=2),
          "delaunay": lambda: G.delaunay(300, seed=3), "er": lambda: G.erdos_renyi_lbc(300, 900, 1000, seed=4)}[case]()
    p = Problem.reduced_graph(*gr)
    ro, cols, w = p.array("row_offsets"), p.array("cols"), p.array("weights")
    got = vertex_features(ro, cols, w)
    want = _features_direct(ro.astype(np.int64), cols.astype(np.int64), w.astype(np.float64))
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_greedy_teacher_is_minimal_and_smaller(seed):
    gr = G.chung_lu_lbc(3000, 2.3, 8.0, 10000, seed=seed)
    p = Problem.reduced_graph(*gr)
    ro, cols = p.array("row_offsets"), p.array("cols")
    rank = fvs_greedy(ro, cols)
    in_f = (rank >= 0).astype(np.uint8)
    assert not induced_has_cycle(ro, cols, in_f)
    for v in np.flatnonzero(in_f)[:200]:  # minimal: dropping any member leaves a cycle
        t = in_f.copy()
        t[v] = 0
        assert induced_has_cycle(ro, cols, t)
    assert int(in_f.sum()) < p.info.fvs_size
    X, y = training_rows(p, max_rows=10 ** 9)
    assert X.shape[0] > p.info.nodes and X.shape[1] == 4 and set(np.unique(y)) <= {0.0, 1.0}


def test_reference_order_is_the_default_pass():
    gr = G.chung_lu_lbc(2000, 2.3, 8.0, 10000, seed=5)
    p = Problem.reduced_graph(*gr)
    ro, cols = p.array("row_offsets"), p.array("cols")
    assert np.array_equal(fvs_reverse_delete(ro, cols, reference_order(ro)), fvs_reverse_delete(ro, cols))
    assert np.array_equal(p.array("fvs_rank"), fvs_reverse_delete(ro, cols))


def test_peel_and_masked_features():
    gr = G.chung_lu_lbc(1500, 2.3, 8.0, 10000, seed=8)
    p = Problem.reduced_graph(*gr)
    ro, cols, w = p.array("row_offsets"), p.array("cols"), p.array("weights")
    n = ro.size - 1
    alive = np.ones(n, dtype=np.uint8)
    alive[np.random.default_rng(1).choice(n, n // 3, replace=False)] = 0
    left = fvs_peel(ro, cols, alive)
    assert left == int(alive.sum())
    # every survivor keeps two edges to survivors; the peeled ones could not lie on a cycle
    ro64, c64 = ro.astype(np.int64), cols.astype(np.int64)
    for v in np.flatnonzero(alive):
        assert alive[c64[ro64[v]:ro64[v + 1]]].sum() >= 2
    # features of the survivors equal the features of the induced subgraph built explicitly
    keep = np.flatnonzero(alive)
    relabel = -np.ones(n, dtype=np.int64)
    relabel[keep] = np.arange(keep.size)
    rows = np.repeat(np.arange(n), np.diff(ro64))
    m = (alive[rows] == 1) & (alive[c64] == 1)
    sr, sc, sw = relabel[rows[m]], relabel[c64[m]], w[m]
    sro = np.concatenate([[0], np.cumsum(np.bincount(sr, minlength=keep.size))]).astype(np.uint32)
    got = vertex_features(ro, cols, w, alive)
    np.testing.assert_allclose(got[keep], vertex_features(sro, sc.astype(np.uint32), sw), rtol=1e-5, atol=1e-6)
    assert not got[alive == 0].any()


@pytest.mark.parametrize("seed", [1, 2])
def test_round_greedy_gives_a_minimal_smaller_fvs(seed):
    gr = G.chung_lu_lbc(4000, 2.3, 8.0, 10000, seed=seed)
    p = Problem.reduced_graph(*gr)
    ro, cols, w = p.array("row_offsets"), p.array("cols"), p.array("weights")
    rank, order = scored_greedy(ro, cols, w, degree_score)
    assert np.array_equal(np.sort(order), np.arange(ro.size - 1))
    in_f = (rank >= 0).astype(np.uint8)
    assert not induced_has_cycle(ro, cols, in_f)
    assert int(in_f.sum()) < p.info.fvs_size
    for v in np.flatnonzero(in_f)[:100]:
        t = in_f.copy()
        t[v] = 0
        assert induced_has_cycle(ro, cols, t)


class _FixedScorer:
    def __init__(self, fn):
        self.fn = fn

    def logits(self, X):
        return self.fn(X)


def test_choose_order_keeps_only_a_smaller_fvs():
    gr = G.chung_lu_lbc(3000, 2.3, 8.0, 10000, seed=6)
    p = Problem.reduced_graph(*gr)
    # remaining degree as the score gives a smaller set than the reference order: used
    good = choose_order(p, _FixedScorer(degree_score))
    assert isinstance(good, LearnedChoice) and good.used and good.order is not None
    assert good.fvs_learned < good.fvs_reference
    assert np.array_equal(fvs_reverse_delete(p.array("row_offsets"), p.array("cols"), good.order) >= 0,
                          scored_greedy(p.array("row_offsets"), p.array("cols"), p.array("weights"),
                                        degree_score)[0] >= 0)
    # taking the lowest degrees first gives a larger set: the reference order is kept
    bad = choose_order(p, _FixedScorer(lambda X: -X[:, 0]))
    assert not bad.used and bad.order is None and bad.fvs_learned >= bad.fvs_reference
