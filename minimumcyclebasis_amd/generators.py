"""Seeded synthetic graph generators for the benchmark and test workloads.

Every generator returns ``(n, u, v, w)``: vertex count and 0-based int32 endpoint/weight arrays
with each undirected edge listed once (u > v), sorted by (u, v).  The same seed gives the same
graph on every machine (numpy PCG64).  ``write_mtx`` writes the symmetric integer
Matrix-Market file the reference reads (include/FileReader.h:27-67).

Workloads (sizes of the configurations the measurements are quoted on):

* ``grid(64, 64, 100, seed=1)``            4096 vertices, 8064 edges, weights U{1..100}
* ``erdos_renyi_lbc(20000, 60000, 1000)``   G(n, m) restricted to its largest biconnected component
* ``delaunay(100000)``                      weight = floor(1000 * length + 0.5)
* ``batched_components(2000, 50, 5000)``    random 2-connected components joined by bridges
* ``chung_lu_lbc(50000, 2.3, 8, 10000)``    power-law expected degrees, largest biconnected component
"""
from __future__ import annotations

import numpy as np


def _finish(n: int, a: np.ndarray, b: np.ndarray, w: np.ndarray):
    hi = np.maximum(a, b).astype(np.int64)
    lo = np.minimum(a, b).astype(np.int64)
    order = np.lexsort((lo, hi))
    return (int(n), hi[order].astype(np.int32), lo[order].astype(np.int32),
            np.asarray(w, dtype=np.int64)[order].astype(np.int32))


def _canonical_pairs(a: np.ndarray, b: np.ndarray):
    """Distinct undirected pairs without self-loops, in (max, min) lexicographic order."""
    hi = np.maximum(a, b).astype(np.int64)
    lo = np.minimum(a, b).astype(np.int64)
    keep = hi != lo
    key = np.unique(hi[keep] * (1 << 32) + lo[keep])
    return (key >> 32).astype(np.int64), (key & 0xFFFFFFFF).astype(np.int64)


def _weights(rng: np.random.Generator, m: int, wmax: int) -> np.ndarray:
    return rng.integers(1, wmax + 1, size=m, dtype=np.int64)


def grid(rows: int, cols: int, wmax: int = 100, seed: int = 1):
    """4-neighbour rows x cols grid; weights i.i.d. U{1..wmax} in (max, min) edge order."""
    idx = np.arange(rows * cols, dtype=np.int64).reshape(rows, cols)
    a = np.concatenate([idx[:, 1:].ravel(), idx[1:, :].ravel()])
    b = np.concatenate([idx[:, :-1].ravel(), idx[:-1, :].ravel()])
    hi, lo = _canonical_pairs(a, b)
    rng = np.random.default_rng(seed)
    return _finish(rows * cols, hi, lo, _weights(rng, hi.size, wmax))


def complete(n: int, wmax: int = 10, seed: int = 1):
    a, b = np.tril_indices(n, -1)
    hi, lo = _canonical_pairs(a, b)
    rng = np.random.default_rng(seed)
    return _finish(n, hi, lo, _weights(rng, hi.size, wmax))


def largest_biconnected_component(n: int, u, v, w):
    """Restrict an edge list to its largest biconnected component (most edges, lowest id on ties);
    vertices are renumbered in increasing original id."""
    from .host import biconnected_components

    u = np.asarray(u, dtype=np.int32)
    v = np.asarray(v, dtype=np.int32)
    w = np.asarray(w, dtype=np.int32)
    comp = biconnected_components(n, u, v)
    if comp.size == 0:
        raise ValueError("graph has no edges")
    counts = np.bincount(comp)
    best = int(np.argmax(counts))
    sel = comp == best
    uu, vv, ww = u[sel], v[sel], w[sel]
    verts = np.unique(np.concatenate([uu, vv]))
    relabel = np.full(n, -1, dtype=np.int64)
    relabel[verts] = np.arange(verts.size)
    return _finish(verts.size, relabel[uu], relabel[vv], ww)


def erdos_renyi_lbc(n: int, m: int, wmax: int = 1000, seed: int = 1):
    """G(n, m) simple graph (m distinct uniform pairs), weights U{1..wmax}, largest biconnected component."""
    rng = np.random.default_rng(seed)
    total = n * (n - 1) // 2
    if m > total:
        raise ValueError("m exceeds the number of vertex pairs")
    chosen = np.empty(0, dtype=np.int64)
    while chosen.size < m:
        a = rng.integers(0, n, size=2 * (m - chosen.size) + 16)
        b = rng.integers(0, n, size=a.size)
        keep = a != b
        key = np.maximum(a, b)[keep] * n + np.minimum(a, b)[keep]
        merged = np.concatenate([chosen, key])
        _, first = np.unique(merged, return_index=True)
        chosen = merged[np.sort(first)][:m]
    hi = chosen // n
    lo = chosen % n
    w = _weights(rng, m, wmax)
    n0, a, b, ww = _finish(n, hi, lo, w)
    return largest_biconnected_component(n0, a, b, ww)


def delaunay(n: int, seed: int = 1, scale: float = 1000.0):
    """Delaunay triangulation of n uniform points in the unit square; weight = floor(scale*len + 0.5)."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    tri = Delaunay(pts)
    s = tri.simplices.astype(np.int64)
    a = np.concatenate([s[:, 0], s[:, 1], s[:, 2]])
    b = np.concatenate([s[:, 1], s[:, 2], s[:, 0]])
    hi, lo = _canonical_pairs(a, b)
    length = np.hypot(*(pts[hi] - pts[lo]).T)
    w = np.floor(scale * length + 0.5).astype(np.int64)
    return _finish(n, hi, lo, w)


def chung_lu_lbc(n: int, gamma: float = 2.3, avg_degree: float = 8.0, wmax: int = 10000, seed: int = 1):
    """Chung-Lu graph with expected degrees k_i proportional to (i+1)^(-1/(gamma-1)), mean avg_degree;
    each pair (i, j) is kept with probability min(1, k_i k_j / sum k), sampled by drawing
    n*avg_degree/2 endpoint pairs proportional to k and dropping loops and duplicates.
    Weights U{1..wmax}; restricted to the largest biconnected component."""
    rng = np.random.default_rng(seed)
    k = (np.arange(n, dtype=np.float64) + 1.0) ** (-1.0 / (gamma - 1.0))
    k *= avg_degree * n / k.sum()
    p = k / k.sum()
    m = int(round(n * avg_degree / 2))
    a = rng.choice(n, size=m, p=p)
    b = rng.choice(n, size=m, p=p)
    hi, lo = _canonical_pairs(a, b)
    w = _weights(rng, hi.size, wmax)
    n0, x, y, ww = _finish(n, hi, lo, w)
    return largest_biconnected_component(n0, x, y, ww)


def random_biconnected(n: int, avg_degree: float = 5.0, wmax: int = 1000, seed: int = 1):
    """A Hamiltonian cycle on a random vertex order plus uniform random chords up to
    n*avg_degree/2 edges (a cycle plus chords is 2-connected)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n).astype(np.int64)
    a = perm
    b = np.roll(perm, -1)
    target = min(max(n, int(round(n * avg_degree / 2))), n * (n - 1) // 2)  # a simple graph has at most n(n-1)/2
    hi, lo = _canonical_pairs(a, b)
    key = hi * n + lo
    while key.size < target:
        x = rng.integers(0, n, size=2 * (target - key.size) + 16)
        y = rng.integers(0, n, size=x.size)
        keep = x != y
        k2 = np.maximum(x, y)[keep] * n + np.minimum(x, y)[keep]
        merged = np.concatenate([key, k2])
        _, first = np.unique(merged, return_index=True)
        key = merged[np.sort(first)][:target]
    hi = key // n
    lo = key % n
    return _finish(n, hi, lo, _weights(rng, key.size, wmax))


def batched_components(k: int = 2000, nmin: int = 50, nmax: int = 5000, deg_lo: float = 4.0,
                       deg_hi: float = 8.0, wmax: int = 1000, seed: int = 1):
    """k random 2-connected components (vertex count log-uniform in [nmin, nmax], average degree
    U[deg_lo, deg_hi]) joined into one connected graph by k-1 bridges (component i to a uniformly
    chosen earlier component).  Returns (n, u, v, w, comp_of_vertex)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=k)).astype(np.int64)
    sizes = np.clip(sizes, max(nmin, 3), nmax)
    us, vs, ws, comp = [], [], [], []
    base = 0
    starts = []
    for i in range(k):
        d = rng.uniform(deg_lo, deg_hi)
        n_i, a, b, w = random_biconnected(int(sizes[i]), d, wmax, seed=int(rng.integers(1 << 31)))
        us.append(a.astype(np.int64) + base)
        vs.append(b.astype(np.int64) + base)
        ws.append(w.astype(np.int64))
        comp.append(np.full(n_i, i, dtype=np.int32))
        starts.append(base)
        base += n_i
    starts = np.asarray(starts)
    bu, bv = [], []
    for i in range(1, k):
        j = int(rng.integers(0, i))
        bu.append(starts[i] + int(rng.integers(0, sizes[i])))
        bv.append(starts[j] + int(rng.integers(0, sizes[j])))
    u = np.concatenate(us + [np.asarray(bu, dtype=np.int64)])
    v = np.concatenate(vs + [np.asarray(bv, dtype=np.int64)])
    w = np.concatenate(ws + [rng.integers(1, wmax + 1, size=len(bu))])
    n, a, b, ww = _finish(base, u, v, w)
    return n, a, b, ww, np.concatenate(comp)


def write_mtx(path: str, n: int, u, v, w) -> None:
    """Symmetric integer coordinate Matrix-Market file, 1-based, one line per undirected edge."""
    u = np.asarray(u, dtype=np.int64) + 1
    v = np.asarray(v, dtype=np.int64) + 1
    w = np.asarray(w, dtype=np.int64)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate integer symmetric\n")
        f.write(f"{n} {n} {u.size}\n")
        body = np.column_stack([u, v, w])
        np.savetxt(f, body, fmt="%d")


def read_mtx(path: str):
    """Edge list of a symmetric integer coordinate Matrix-Market file (0-based)."""
    with open(path) as f:
        banner = f.readline()
        if not banner.lower().startswith("%%matrixmarket"):
            raise ValueError("not a Matrix-Market file")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n, _, m = (int(x) for x in line.split())
        data = np.loadtxt(f, dtype=np.int64, ndmin=2) if m else np.zeros((0, 3), dtype=np.int64)
    return n, (data[:, 0] - 1).astype(np.int32), (data[:, 1] - 1).astype(np.int32), data[:, 2].astype(np.int32)
