"""The native round state of the learned FVS construction (host.FvsRounds) against the construction
restated round by round over the whole graph: features of the alive subgraph (host.vertex_features with
an alive mask), a stable sort of the negated scores (ties to the lower id, NaN last), the picks cleared
and the rest peeled (host.fvs_peel).  Every round's ids, feature rows and picks must be identical, on
graphs with many ties (degree scores), with random and NaN scores, and with picks that peel each other."""
import numpy as np
import pytest

from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.host import FvsRounds, Problem, fvs_peel, vertex_features
from minimumcyclebasis_amd.learned_fvs import degree_score, graph_of, scored_greedy


def rounds_by_full_scans(ro, cols, w, score_fn, frac=0.02):
    ro = np.ascontiguousarray(ro, dtype=np.uint32)
    n = ro.size - 1
    alive = np.ones(n, dtype=np.uint8)
    left = fvs_peel(ro, cols, alive)
    out = []
    while left > 0:
        ids = np.flatnonzero(alive).astype(np.int32)
        X = vertex_features(ro, cols, w, alive)[ids]
        s = np.asarray(score_fn(ids, X), dtype=np.float32)
        k = max(1, int(np.ceil(frac * ids.size)))
        top = ids[np.argsort(-s, kind="stable")[:k]]
        out.append((ids, X, top))
        alive[top] = 0
        left = fvs_peel(ro, cols, alive)
    return out


def rounds_native(ro, cols, w, score_fn, frac=0.02):
    r = FvsRounds(ro, cols, w)
    out = []
    while r.left > 0:
        ids, X = r.features()
        s = np.asarray(score_fn(ids, X), dtype=np.float32)
        top = r.take_top(s, max(1, int(np.ceil(frac * ids.size))))
        out.append((ids, X, top))
    r.close()
    return out


def _degree(ids, X):
    return X[:, 0]


def _hashed(ids, X):
    return (np.sin(ids.astype(np.float64) * 12.9898 + X[:, 1] * 78.233) * 43758.5453 % 1.0).astype(np.float32)


def _with_nans(ids, X):
    s = X[:, 0] - X[:, 2]
    s[ids % 5 == 0] = np.nan
    s[ids % 7 == 0] = -np.inf
    return s


GRAPHS = {
    "delaunay": lambda: G.delaunay(3000, seed=3),
    "chunglu": lambda: G.chung_lu_lbc(4000, 2.3, 8.0, 100, seed=2),
    "er_multi": lambda: G.erdos_renyi_lbc(1500, 4000, 5, seed=7),
    "grid": lambda: G.grid(40, 30, 3, seed=1),
    "biconn": lambda: G.random_biconnected(800, 2.4, 4, seed=5),
}


@pytest.mark.parametrize("score", [_degree, _hashed, _with_nans], ids=["degree", "hashed", "nan"])
@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("frac", [0.02, 0.3])
def test_native_rounds_equal_full_scans(name, score, frac):
    p = Problem.reduced_graph(*GRAPHS[name]())
    ro, cols, w = graph_of(p)
    want = rounds_by_full_scans(ro, cols, w, score, frac)
    got = rounds_native(ro, cols, w, score, frac)
    assert len(got) == len(want) and len(want) > 1
    for i, ((a, x, t), (b, y, u)) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), f"round {i}: ids"
        assert np.array_equal(x, y), f"round {i}: features"
        assert np.array_equal(t, u), f"round {i}: picks"
    p.close()


def test_scored_greedy_takes_the_native_rounds():
    """scored_greedy's picks are the rounds' picks, passed to reverse-delete last to first."""
    p = Problem.reduced_graph(*G.chung_lu_lbc(3000, 2.3, 8.0, 100, seed=4))
    ro, cols, w = graph_of(p)
    seen = []
    rank, order = scored_greedy(ro, cols, w, degree_score, on_round=lambda ids, X: seen.append(ids.size))
    want = rounds_by_full_scans(ro, cols, w, _degree)
    picks = np.concatenate([t for _, _, t in want])
    assert seen == [ids.size for ids, _, _ in want]
    assert np.array_equal(order[-picks.size:], picks[::-1])
    assert np.array_equal(np.sort(order), np.arange(ro.size - 1))
    assert (rank >= 0).sum() <= picks.size
    p.close()


def test_take_top_checks_its_inputs():
    p = Problem.reduced_graph(*G.grid(8, 8, 2, seed=1))
    ro, cols, w = graph_of(p)
    r = FvsRounds(ro, cols, w)
    with pytest.raises(RuntimeError):
        r.take_top(np.zeros(3, np.float32), 1)  # no feature rows read yet
    ids, X = r.features()
    with pytest.raises(RuntimeError):
        r.take_top(np.zeros(ids.size + 1, np.float32), 1)
    top = r.take_top(X[:, 0], ids.size + 5)  # k above the row count takes every row
    assert top.size == ids.size and r.left == 0
    r.close()
    p.close()
