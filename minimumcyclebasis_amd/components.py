"""Biconnected-component split of an input graph, in process.

The reference's pipeline (main.sh:1-30) runs ``bicc/bicc_decomposition`` to write every biconnected component to a
file, relabels each (``bicc/Relabeller``) and runs the MCB driver once per component, appending one stats row per
component.  The minimum cycle basis of a graph is the union of its components' bases, since every cycle lies in one
biconnected component.  Here the split is ``host.biconnected_components`` (iterative Tarjan in libmcb_host.so),
and each component's vertices are renumbered 0.. in ascending order of their input ids.  The components' weights,
counts and cycle-space dimensions are exact.  Under weight ties, the particular cycles chosen inside a component may
differ from main.sh's: those depend on the relabelling, and this split does not reproduce the order in which the
bicc tool writes a component's edges.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .host import biconnected_components


@dataclass
class Component:
    index: int
    n: int
    u: np.ndarray        # local endpoints, 0-based
    v: np.ndarray
    w: np.ndarray
    vertices: np.ndarray  # input id of each local vertex

    @property
    def cycle_space_dim(self) -> int:
        return int(self.u.size) - self.n + 1


def build_problems(comps: list[Component], threads: int = 1, **options) -> list:
    """Host preprocessing (``host.Problem.from_edges``) of every component, ``threads`` components at a
    time, each on one thread (the library call releases the interpreter lock).  Results in input order."""
    from concurrent.futures import ThreadPoolExecutor

    from .host import Problem

    options.setdefault("num_threads", 1)

    def one(c):
        return Problem.from_edges(c.n, c.u, c.v, c.w, **options)

    if threads <= 1:
        return [one(c) for c in comps]
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, comps))


def split_components(n: int, u, v, w, min_edges: int = 2) -> list[Component]:
    """Biconnected components of a simple undirected graph with at least ``min_edges`` edges (a single edge, a
    bridge, has no cycle), in the order the decomposition completes them."""
    u = np.ascontiguousarray(u, dtype=np.int64)
    v = np.ascontiguousarray(v, dtype=np.int64)
    w = np.ascontiguousarray(w, dtype=np.int64)
    comp = biconnected_components(n, u, v)
    out = []
    if comp.size == 0:
        return out
    order = np.argsort(comp, kind="stable")
    bounds = np.flatnonzero(np.diff(comp[order])) + 1
    for idx in np.split(order, bounds):
        if idx.size < min_edges:
            continue
        cu, cv, cw = u[idx], v[idx], w[idx]
        verts = np.unique(np.concatenate([cu, cv]))
        lu = np.searchsorted(verts, cu)
        lv = np.searchsorted(verts, cv)
        out.append(Component(int(comp[idx[0]]), int(verts.size), lu.astype(np.int32), lv.astype(np.int32),
                             cw.astype(np.int32), verts))
    return out
