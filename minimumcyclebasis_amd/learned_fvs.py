"""Learned FVS construction: a small MLP scores the vertices of the reduced multigraph, and the greedy
FVS construction takes them in score order.

The reference builds its feedback vertex set by reverse-delete over a static degree order
(include/FVS.cpp:166-227, restated in libmcb_host.so).  Fewer FVS vertices mean fewer shortest-path
roots, so fewer tree positions and candidate cycles for the whole phase loop.  Here a 4 -> 64 -> 64 -> 1
MLP (``include/mcb_mlp.h``; BF16 weights, FP32 accumulation, forward and backward kernels on the GPU)
scores vertices from four local features of the graph that is still left (``host.vertex_features``:
degree, neighbours' degree sum, core number, mean incident weight).  The construction works in rounds:
vertices that lie on no remaining cycle are peeled away (``host.fvs_peel``), the remaining ones are
scored, and the best-scored 2 % are taken; then the picks are made a minimal FVS by the reverse-delete
pass (the other vertices first in the reference order, then the picks from the last to the first).
The rounds run in one native state (``host.FvsRounds``) that compacts the graph as it shrinks, so the
whole construction costs time in the sum of the round sizes rather than rounds x graph size.
The MLP is trained to predict membership in the set that the one-at-a-time peeling max-degree greedy
finds (``host.fvs_greedy``), on rows of that process's own rounds, over graphs of the same generator.

Safety: the pass returns a minimal FVS whatever the picks; the result is also checked acyclic
(``host.induced_has_cycle``), and the reference order is used whenever the learned set is not smaller.
The MCB weight does not depend on the root set (it is unique); the chosen cycles may, under ties.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from .host import FvsRounds, Problem, fvs_greedy, fvs_reverse_delete, induced_has_cycle

N_PARAMS = 4545
PARAMS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "fvs_mlp.json")


def _check(rc: int) -> int:
    if rc < 0:
        raise _lib.MCBError(rc, _lib.hip().mcb_mlp_last_error().decode(errors="replace"))
    return rc


class Scorer:
    """The device MLP (``mcb_mlp_*``)."""

    def __init__(self, params: np.ndarray | None = None):
        h = ctypes.c_void_p()
        _check(_lib.hip().mcb_mlp_create(ctypes.byref(h)))
        self._h = h
        if params is not None:
            self.set_params(params)

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters")
        _check(_lib.hip().mcb_mlp_set_params(self._h, p.ctypes.data))

    def params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, dtype=np.float32)
        _check(_lib.hip().mcb_mlp_get_params(self._h, out.ctypes.data))
        return out

    def logits(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(X.shape[0], dtype=np.float32)
        _check(_lib.hip().mcb_mlp_forward(self._h, X.ctypes.data, X.shape[0], out.ctypes.data))
        return out

    def gradient(self, X: np.ndarray, y: np.ndarray) -> tuple[float, np.ndarray]:
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        _check(_lib.hip().mcb_mlp_gradient(self._h, X.ctypes.data, y.ctypes.data, X.shape[0], g.ctypes.data,
                                           ctypes.byref(loss)))
        return loss.value, g

    def fit(self, X: np.ndarray, y: np.ndarray, epochs: int = 8, batch: int = 4096, lr: float = 3e-3,
            seed: int = 0, log=None) -> list[float]:
        """Adam on the device over (X, y), a fresh random row order each epoch; returns the epoch losses."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        _check(_lib.hip().mcb_mlp_set_data(self._h, X.ctypes.data, y.ctypes.data, X.shape[0]))
        rng = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            order = rng.permutation(X.shape[0]).astype(np.int32)
            loss = ctypes.c_float()
            _check(_lib.hip().mcb_mlp_train_epoch(self._h, order.ctypes.data, int(batch), float(lr),
                                                  ctypes.byref(loss)))
            losses.append(loss.value)
            if log:
                log(f"epoch {ep}: loss {loss.value:.5f}")
        return losses

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.hip().mcb_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_params(path: str = PARAMS_FILE) -> np.ndarray:
    with open(path) as f:
        d = json.load(f)
    p = np.asarray(d["params"], dtype=np.float32)
    if p.size != N_PARAMS:
        raise ValueError(f"{path}: expected {N_PARAMS} parameters")
    return p


def save_params(params: np.ndarray, path: str, meta: dict) -> None:
    with open(path, "w") as f:
        json.dump({**meta, "params": [float(x) for x in np.asarray(params, dtype=np.float32)]}, f)


def graph_of(problem: Problem):
    """(row offsets, cols, weights) of a problem's reduced multigraph."""
    return problem.array("row_offsets"), problem.array("cols"), problem.array("weights")


def reference_order(row_offsets) -> np.ndarray:
    """The reference's reverse-delete order: degree ascending, then id descending (FVS.cpp:172-193)."""
    ro = np.asarray(row_offsets, dtype=np.int64)
    n = ro.size - 1
    deg = np.diff(ro)
    return np.lexsort((-np.arange(n), deg)).astype(np.int32)


ROUND_FRACTION = 0.02  # share of the remaining vertices taken per round


def scored_greedy(row_offsets, cols, weights, score_fn, frac: float = ROUND_FRACTION, on_round=None):
    """Round-based greedy FVS: peel, score the remaining vertices (``score_fn`` on their feature rows, higher
    = take first), take the best ceil(frac * remaining) (ties: lower id), repeat; then minimal by
    reverse-delete.  Returns (root rank per vertex, the order passed to the pass).  ``on_round(ids, X)``
    sees every round's remaining vertices and their features."""
    ro = np.ascontiguousarray(row_offsets, dtype=np.uint32)
    n = ro.size - 1
    rounds = FvsRounds(ro, cols, weights)
    picks = []
    try:
        while rounds.left > 0:
            ids, X = rounds.features()
            if on_round:
                on_round(ids, X)
            s = np.asarray(score_fn(X), dtype=np.float32)
            picks.append(rounds.take_top(s, max(1, int(np.ceil(frac * ids.size)))))
    finally:
        rounds.close()
    picks = np.concatenate(picks) if picks else np.zeros(0, dtype=np.int64)
    in_pick = np.zeros(n, dtype=bool)
    in_pick[picks] = True
    base = reference_order(ro)
    order = np.concatenate([base[~in_pick[base]], picks[::-1]]).astype(np.int32)
    return fvs_reverse_delete(ro, cols, order), order


def degree_score(X: np.ndarray) -> np.ndarray:
    """The scorer-free round greedy: the remaining degree alone (feature 0)."""
    return X[:, 0]


def training_rows(problem: Problem, max_rows: int = 20000, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Rows of the rounds of the degree-driven round greedy on a reduced graph (``Problem.reduced_graph``),
    labelled with membership in the one-at-a-time greedy's FVS; at most max_rows, drawn uniformly."""
    ro, cols, w = graph_of(problem)
    label = (fvs_greedy(ro, cols) >= 0).astype(np.float32)
    xs, ys = [], []

    def keep(ids, X):
        xs.append(X)
        ys.append(label[ids])

    scored_greedy(ro, cols, w, degree_score, on_round=keep)
    X, y = np.concatenate(xs), np.concatenate(ys)
    if X.shape[0] > max_rows:
        pick = np.sort(np.random.default_rng(seed).choice(X.shape[0], max_rows, replace=False))
        X, y = X[pick], y[pick]
    return X, y


@dataclass
class LearnedChoice:
    order: np.ndarray | None  # the order to pass as fvs_order, or None for the reference order
    fvs_reference: int
    fvs_learned: int
    used: bool


def choose_order(problem: Problem, scorer) -> LearnedChoice:
    """The learned construction on a reduced graph, kept only if its FVS is verified and smaller."""
    ro, cols, w = graph_of(problem)
    rank, order = scored_greedy(ro, cols, w, scorer.logits)
    f_learned = int((rank >= 0).sum())
    f_ref = int(problem.info.fvs_size)
    ok = f_learned < f_ref and not induced_has_cycle(ro, cols, (rank >= 0).astype(np.uint8))
    return LearnedChoice(order if ok else None, f_ref, f_learned, ok)


_default_scorer = None


def default_scorer() -> Scorer:
    """A scorer with the shipped parameters (``data/fvs_mlp.json``), created once."""
    global _default_scorer
    if _default_scorer is None:
        _default_scorer = Scorer(load_params())
    return _default_scorer


def learned_order(n: int, u, v, w, reduce_degree_two: bool = True, scorer: Scorer | None = None) -> LearnedChoice:
    """Reduce the graph, score its vertices and return the order for ``fvs_order`` (None: reference order)."""
    g = Problem.reduced_graph(n, u, v, w, reduce_degree_two=reduce_degree_two)
    try:
        if g.info.single_cycle:
            return LearnedChoice(None, 1, 1, False)
        return choose_order(g, scorer or default_scorer())
    finally:
        g.close()


def training_set(seeds, nmin: int = 2000, nmax: int = 20000, log=None):
    """Rows of Chung-Lu graphs (gamma 2.3, average degree 8, largest biconnected component, weights
    U{1..10000}) with n drawn log-uniformly from [nmin, nmax] per seed."""
    from . import generators as G

    X, Y = [], []
    for i, seed in enumerate(seeds):
        r = np.random.default_rng(10_000 + int(seed))
        n = int(np.exp(r.uniform(np.log(nmin), np.log(nmax))))
        gr = G.chung_lu_lbc(n, 2.3, 8.0, 10000, seed=int(seed))
        p = Problem.reduced_graph(*gr)
        try:
            if p.info.single_cycle:
                continue
            x, y = training_rows(p, seed=int(seed))
        finally:
            p.close()
        X.append(x)
        Y.append(y)
        if log and (i + 1) % 50 == 0:
            log(f"{i + 1} graphs, {sum(len(a) for a in Y)} rows")
    return np.concatenate(X), np.concatenate(Y)


def main(argv=None) -> int:
    """python -m minimumcyclebasis_amd.learned_fvs train --out PATH [--graphs 500] [--epochs 8]"""
    import argparse
    import sys
    import time

    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--out", default=PARAMS_FILE)
    ap.add_argument("--graphs", type=int, default=500)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)

    def log(msg):
        print(f"[train] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    X, Y = training_set(range(a.graphs), log=log)
    log(f"training set: {X.shape[0]} rows, {Y.mean():.3f} positive, {time.perf_counter() - t0:.1f} s")
    rng = np.random.default_rng(a.seed)
    # He-style initialisation for the ReLU layers, small output layer
    init = np.concatenate([rng.normal(0, np.sqrt(2 / 4), 256), np.zeros(64), rng.normal(0, np.sqrt(2 / 64), 4096),
                           np.zeros(64), rng.normal(0, np.sqrt(1 / 64), 64), [0.0]]).astype(np.float32)
    sc = Scorer(init)
    t1 = time.perf_counter()
    losses = sc.fit(X, Y, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
    train_s = time.perf_counter() - t1
    params = sc.params()
    save_params(params, a.out, {"graphs": a.graphs, "rows": int(X.shape[0]), "epochs": a.epochs, "batch": a.batch,
                                "lr": a.lr, "seed": a.seed, "losses": losses, "train_s": round(train_s, 3),
                                "generator": "chung_lu_lbc(n, 2.3, 8.0, 10000, seed), n log-uniform in [2000, 20000]"})
    log(f"trained in {train_s:.2f} s, saved {a.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
