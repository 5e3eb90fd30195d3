"""Pins the oracle's per-phase checks on the position layout (oracle_phase_select, oracle_cycle_vector,
oracle_orthogonality), which check the full-size graphs phase by phase:

* against the reference's own traces: given the reference's printed S_e, the scan picks the reference's
  chosen cycle and the walk gives its printed C_e (tiny graphs, mcb/Test.cpp:239-300);
* against the whole-loop oracle (oracle_run, itself pinned to the reference's chosen-cycle sequences):
  given its S_e, every phase's selection and cycle vector, on pruned and unpruned layouts;
* the orthogonality combinations against a direct computation on random vectors, and planted errors
  (a flipped support bit, a wrong selection) that each check must catch."""
import numpy as np
import pytest

from helpers import load_trace, problem_of, trace_names, words_of
from oracle import pyoracle

MODES = ("reduced", "plain")


def dot(a: np.ndarray, b: np.ndarray) -> int:
    x = np.bitwise_and(a, b)
    return int(sum(int(t).bit_count() for t in x) & 1)


@pytest.mark.parametrize("prune", [False, True])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", trace_names())
def test_phase_select_and_cycle_follow_the_loop(name, mode, prune):
    fx = load_trace(name)
    p = problem_of(fx, mode, keep_vertex_trees=True, prune_trees=prune)
    res = pyoracle.run(p, num_threads=2, want_s=True, want_c=True)
    sel, S, C = res["selected"], res["S"], res["C"]
    N = p.info.num_non_tree
    # every phase of the tiny graphs; the first 16 and 24 spread over the loop of the medium ones
    phases = range(N) if fx["tier"] == "tiny" else sorted(set(range(min(N, 16))) | set(range(0, N, max(1, N // 24))))
    for e in phases:
        assert pyoracle.phase_select(p, S[e], taken=sel[:e]) == sel[e], f"phase {e}"
        # erasing chosen candidates changes nothing: an earlier winner is even for every later S_e
        assert pyoracle.phase_select(p, S[e]) == sel[e], f"phase {e} without erasure"
        assert np.array_equal(pyoracle.cycle_vector(p, int(sel[e])), C[e]), f"phase {e} cycle vector"
    out = pyoracle.orthogonality(p, sel, np.arange(N), S, seed=3)
    assert not out.any()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", trace_names("tiny"))
def test_phase_select_on_the_reference_vectors(name, mode):
    """Given the reference's printed S_e, the scan picks the reference's chosen cycle, and the walk gives
    the reference's printed C_e."""
    fx = load_trace(name)
    ref = fx[mode]
    p = problem_of(fx, mode, prune_trees=True)
    W = p.info.words
    r, a, b, w = p.candidate_table()
    taken = []
    for e, ph in enumerate(ref["phases"]):
        k = pyoracle.phase_select(p, words_of(ph["S"], W), taken=taken)
        assert k >= 0
        assert [int(r[k]) + 1, int(a[k]) + 1, int(b[k]) + 1, int(w[k])] == ph["chosen"], f"phase {e}"
        assert np.array_equal(pyoracle.cycle_vector(p, k), words_of(ph["C"], W)), f"phase {e}"
        taken.append(k)


def test_phase_select_without_odd_candidate():
    fx = load_trace("grid4x4")
    p = problem_of(fx, "reduced", prune_trees=True)
    assert pyoracle.phase_select(p, np.zeros(p.info.words, dtype=np.uint64)) == -1


@pytest.mark.parametrize("name", ["grid6x5", "petersen", "er120_ties", "delaunay150"])
def test_orthogonality_combinations_match_direct_products(name):
    """Bit t of the result is XOR over i < j with bit t of r[i] of <C_i, S_j>, on random (non-orthogonal) S."""
    fx = load_trace(name)
    p = problem_of(fx, "reduced", keep_vertex_trees=True, prune_trees=True)
    sel = pyoracle.run(p, num_threads=2)["selected"]
    N, W = p.info.num_non_tree, p.info.words
    rng = np.random.default_rng(11)
    samp = np.unique(rng.integers(0, N, size=min(N, 12)))
    Ss = rng.integers(0, 2**64, size=(samp.size, W), dtype=np.uint64)
    if N % 64:
        Ss[:, -1] &= np.uint64((1 << (N % 64)) - 1)  # coordinates are < N
    out = pyoracle.orthogonality(p, sel, samp, Ss, seed=7)
    r = np.random.default_rng(7).integers(0, 2**64, size=N, dtype=np.uint64)
    Cs = np.stack([pyoracle.cycle_vector(p, int(k)) for k in sel])
    for s, j in enumerate(samp):
        want = np.uint64(0)
        for i in range(j):
            if dot(Cs[i], Ss[s]):
                want ^= r[i]
        assert out[s] == want, f"sample {j}"


@pytest.mark.parametrize("name", ["grid6x5", "er120_ties", "chunglu300"])
def test_planted_errors_are_caught(name):
    fx = load_trace(name)
    p = problem_of(fx, "reduced", keep_vertex_trees=True, prune_trees=True)
    res = pyoracle.run(p, num_threads=2, want_s=True, want_c=True)
    sel, S, C = res["selected"], res["S"], res["C"]
    N = p.info.num_non_tree
    j = N - 1
    # a support bit flipped at a coordinate some earlier cycle holds breaks orthogonality
    c = int(np.flatnonzero(np.unpackbits(C[0].view(np.uint8), bitorder="little"))[0])
    bad = S[j].copy()
    bad[c // 64] ^= np.uint64(1) << np.uint64(c % 64)
    assert pyoracle.orthogonality(p, sel, [j], bad[None, :], seed=1)[0] != 0
    # S_j checked against a wrong earlier selection (the same cycle twice) fails too
    sel2 = sel.copy()
    sel2[1] = sel2[0]
    assert pyoracle.orthogonality(p, sel2, [j], S[j][None, :], seed=1)[0] != 0 or dot(C[1], S[j]) == dot(C[0], S[j])
    # a support vector with its phase's bit cleared picks another candidate (or none)
    e = N // 2
    wrong = S[e].copy()
    wrong[e // 64] ^= np.uint64(1) << np.uint64(e % 64)
    assert pyoracle.phase_select(p, wrong, taken=sel[:e]) != sel[e]


@pytest.mark.parametrize("name", ["grid6x5", "er120_ties", "delaunay150"])
def test_timed_phase_is_one_phase_of_the_loop(name):
    """The CPU baseline's timed phase, started from the loop's state after e phases, chooses the loop's
    rank and leaves the later vectors in the loop's state after e + 1 phases."""
    fx = load_trace(name)
    p = problem_of(fx, "reduced", keep_vertex_trees=True, prune_trees=True)
    N, F, n = p.info.num_non_tree, p.info.fvs_size, p.info.nodes
    sel = pyoracle.run(p, num_threads=2)["selected"]
    for e in sorted({0, 1, N // 3, N - 2}):
        before = pyoracle.run(p, phases=e, want_s=True, num_threads=2)["S"]
        after = pyoracle.run(p, phases=e + 1, want_s=True, num_threads=2)["S"]
        P = np.ascontiguousarray(pyoracle.parity(p, before[e]).reshape(-1))
        later = np.ascontiguousarray(before[e + 1:])
        k, *_ = pyoracle.phase_timed(p, e, before[e], sel[:e], P, later, np.arange(F, dtype=np.int32),
                                     num_threads=2)
        assert k == sel[e]
        assert np.array_equal(later, after[e + 1:]), f"phase {e}: update of the later vectors"
