"""Checks of a whole phase loop at sizes the whole-loop oracle cannot run (no F x n vertex rows), through the
oracle's per-phase checks on the position layout:

* the selection: N distinct valid ranks, N = m' - n' + 1;
* on sampled phases j, given the GPU's own S_j: S_j has bit j and no higher bit, the GPU's choice is the
  first odd candidate in rank order (the reference's list scan, mcb/mcb_cycle.cpp:254-285), and
  <C_j, S_j> = 1 for the cycle vector the oracle builds from the chosen rank;
* <C_i, S_j> = 0 for every i < j at each sample, through 64 random combinations of the earlier cycles.

With the selections before j fixed, S_j is the only vector with bit j, no higher bit and zero products with
C_0 .. C_{j-1}, so the samples pin each S_j exactly and the selection at j given it."""
from __future__ import annotations

import numpy as np

from oracle import pyoracle


def sample_phases(N: int, first: int = 10, spread: int = 50, last: int = 4) -> np.ndarray:
    s = set(range(min(first, N))) | set(np.linspace(0, N - 1, spread).astype(int).tolist())
    s |= set(range(max(0, N - last), N))
    return np.array(sorted(s), dtype=np.int32)


def dot(a: np.ndarray, b: np.ndarray) -> int:
    x = np.bitwise_and(a, b)
    return int(sum(int(t).bit_count() for t in x) & 1)


def check_loop(problem, sel: np.ndarray, rows_of, samples: np.ndarray, seed: int = 0, label: str = "") -> None:
    """sel: the GPU's N selections; rows_of(samples) -> their support vectors (len(samples) x W)."""
    info = problem.info
    N, K = info.num_non_tree, info.num_candidates
    assert N == info.directed_edges // 2 - info.nodes + 1, "cycle-space dimension m' - n' + 1"
    assert sel.size == N and sel.min() >= 0 and sel.max() < K, label
    assert np.unique(sel).size == N, f"{label}: a rank chosen twice"
    rows = rows_of(samples)
    for s, j in enumerate(samples):
        bits = np.flatnonzero(np.unpackbits(rows[s].view(np.uint8), bitorder="little"))
        assert bits.max() == j, f"{label} phase {j}: S_j must hold bit j and no higher bit"
        k = pyoracle.phase_select(problem, rows[s], taken=sel[:j])
        assert k == sel[j], f"{label} phase {j}: GPU chose rank {sel[j]}, the list scan {k}"
        assert dot(pyoracle.cycle_vector(problem, int(sel[j])), rows[s]) == 1, f"{label} phase {j}: <C_j, S_j>"
    out = pyoracle.orthogonality(problem, sel, samples, rows, seed=seed)
    bad = np.flatnonzero(out)
    assert bad.size == 0, f"{label}: S_j not orthogonal to earlier cycles at phases {samples[bad][:8].tolist()}"
