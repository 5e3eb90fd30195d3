"""GPU tests of the learned FVS scorer: the device MLP's forward pass and gradient against the CPU oracle
(oracle/mlp.py with the device's BF16 rounding points), training on the device lowering the loss, and the
shipped parameters giving smaller feedback vertex sets than the reference order on held-out graphs of the
training generator, and smaller ones than the same round construction scored by degree alone (what the
network adds), with the same MCB weight."""
import numpy as np
import pytest

from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.device import minimum_cycle_basis
from minimumcyclebasis_amd.host import Problem
from minimumcyclebasis_amd.learned_fvs import (N_PARAMS, Scorer, choose_order, default_scorer, degree_score,
                                               graph_of, scored_greedy, training_rows, training_set)
from oracle import mlp

pytestmark = pytest.mark.gpu


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, N_PARAMS).astype(np.float32)


def _rows(seed, n=3000):
    p = Problem.reduced_graph(*G.chung_lu_lbc(n, 2.3, 8.0, 10000, seed=seed))
    X, y = training_rows(p)
    p.close()
    return X, y


@pytest.mark.parametrize("rows", [1, 15, 16, 17, 1000, 40000])
def test_forward_matches_oracle(rows):
    X, _ = _rows(1, 20000)
    X = X[:rows] if rows <= X.shape[0] else np.tile(X, (rows // X.shape[0] + 1, 1))[:rows]
    p = _params(2)
    print(f"case forward rows={rows}", flush=True)
    sc = Scorer(p)
    got = sc.logits(X)
    want = mlp.forward(p, X, emulate_bf16=True)
    np.testing.assert_allclose(got, want, rtol=2e-3, atol=2e-3 * (1 + np.abs(want).max()))
    sc.close()


@pytest.mark.parametrize("rows", [16, 37, 5000])
def test_gradient_matches_oracle(rows):
    X, y = _rows(3)
    X, y = X[:rows], y[:rows]
    p = _params(4)
    print(f"case gradient rows={rows}", flush=True)
    sc = Scorer(p)
    loss, g = sc.gradient(X, y)
    want_loss, want = mlp.loss_and_grad(p, X, y, emulate_bf16=True)
    assert abs(loss - want_loss) <= 1e-4 * (1 + abs(want_loss))
    scale = np.abs(want).max()
    for name, lo, hi in (("W1", 0, 256), ("b1", 256, 320), ("W2", 320, 4416), ("b2", 4416, 4480),
                         ("W3", 4480, 4544), ("b3", 4544, 4545)):
        err = np.abs(g[lo:hi] - want[lo:hi]).max()
        assert err <= 1e-2 * scale, (name, err, scale)
    # the parameters are unchanged by a gradient evaluation
    assert np.array_equal(sc.params(), p)
    sc.close()


def test_training_lowers_the_loss():
    X, y = training_set(range(20), nmin=1000, nmax=4000)
    print(f"case train rows={X.shape[0]}", flush=True)
    rng = np.random.default_rng(0)
    init = np.concatenate([rng.normal(0, 0.7, 256), np.zeros(64), rng.normal(0, 0.18, 4096), np.zeros(64),
                           rng.normal(0, 0.12, 64), [0.0]]).astype(np.float32)
    sc = Scorer(init)
    loss0, _ = sc.gradient(X, y)
    losses = sc.fit(X, y, epochs=4, batch=2048, lr=3e-3)
    loss1, _ = sc.gradient(X, y)
    assert losses[-1] < losses[0] and loss1 < 0.9 * loss0, (loss0, losses, loss1)
    # the device's forward pass agrees with the oracle on the trained parameters
    p = sc.params()
    want = mlp.forward(p, X[:2000])
    np.testing.assert_allclose(sc.logits(X[:2000]), want, rtol=2e-3, atol=2e-3 * (1 + np.abs(want).max()))
    sc.close()


def test_shipped_scorer_shrinks_the_fvs():
    """Held-out seeds of the training generator: the learned order's FVS is smaller in total than the
    reference order's, and never used when it is not smaller."""
    sc = default_scorer()
    ref = learned = used = 0
    for seed in range(900, 906):
        n = [2500, 5000, 8000, 12000, 16000, 20000][seed - 900]
        p = Problem.reduced_graph(*G.chung_lu_lbc(n, 2.3, 8.0, 10000, seed=seed))
        ch = choose_order(p, sc)
        print(f"case held-out seed={seed} n={n} F ref={ch.fvs_reference} learned={ch.fvs_learned}", flush=True)
        ref += ch.fvs_reference
        learned += ch.fvs_learned if ch.used else ch.fvs_reference
        used += ch.used
        assert ch.used == (ch.fvs_learned < ch.fvs_reference)
        p.close()
    assert used >= 1 and learned < ref, (learned, ref)


def test_trained_network_beats_degree_scoring():
    """What the network adds: on held-out seeds of the training generator (n log-uniform in [2000, 20000]),
    the round construction scored by the trained network takes fewer roots in total than the same
    construction scored by the remaining degree alone (learned_fvs.degree_score)."""
    sc = default_scorer()
    net = deg = 0
    for seed in range(900, 912):
        n = int(np.exp(np.random.default_rng(seed).uniform(np.log(2000), np.log(20000))))
        p = Problem.reduced_graph(*G.chung_lu_lbc(n, 2.3, 8.0, 10000, seed=seed))
        ro, cols, w = graph_of(p)
        a = int((scored_greedy(ro, cols, w, sc.logits)[0] >= 0).sum())
        b = int((scored_greedy(ro, cols, w, degree_score)[0] >= 0).sum())
        print(f"case held-out seed={seed} n={n} F network={a} degree={b} reference={p.info.fvs_size}", flush=True)
        net += a
        deg += b
        p.close()
    assert net < deg, (net, deg)


@pytest.mark.parametrize("seed", [11, 12])
def test_learned_fvs_keeps_the_weight(seed):
    gr = G.chung_lu_lbc(2500, 2.3, 8.0, 10000, seed=seed)
    print(f"case weight seed={seed}", flush=True)
    plain = minimum_cycle_basis(gr)
    learned = minimum_cycle_basis(gr, fvs="learned")
    assert learned.total_weight == plain.total_weight
    assert learned.num_cycles == plain.num_cycles
    assert learned.stats.fvs_size <= plain.stats.fvs_size
