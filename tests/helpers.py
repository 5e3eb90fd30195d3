"""Shared test helpers: golden fixtures written by tests/golden/make_golden.py and problems built
from them."""
from __future__ import annotations

import glob
import gzip
import hashlib
import json
import os

import numpy as np

from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.host import Problem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def trace_names(tier: str | None = None):
    names = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, "trace_*.json.gz"))):
        name = os.path.basename(p)[len("trace_"):-len(".json.gz")]
        if tier is None or load_trace(name)["tier"] == tier:
            names.append(name)
    return names


_cache: dict = {}


def load_trace(name: str) -> dict:
    if name not in _cache:
        with gzip.open(os.path.join(GOLDEN, f"trace_{name}.json.gz"), "rt") as f:
            _cache[name] = json.load(f)
    return _cache[name]


def csv_names():
    return sorted(os.path.basename(p)[4:-5] for p in glob.glob(os.path.join(GOLDEN, "csv_*.json")))


def load_csv(name: str) -> dict:
    with open(os.path.join(GOLDEN, f"csv_{name}.json")) as f:
        return json.load(f)


def edge_digest(n, u, v, w) -> str:
    h = hashlib.sha256()
    h.update(np.asarray([n], dtype=np.int64).tobytes())
    for a in (u, v, w):
        h.update(np.ascontiguousarray(a, dtype=np.int64).tobytes())
    return h.hexdigest()


def make_graph(gen: str, kw: dict):
    if gen == "explicit":
        e = np.asarray(kw["edges"], dtype=np.int64)
        return G._finish(kw["n"], e[:, 0], e[:, 1], e[:, 2])
    return getattr(G, gen)(**kw)


def graph_of(fx: dict):
    n, u, v, w = make_graph(fx["generator"], fx["kwargs"])
    assert edge_digest(n, u, v, w) == fx["edge_digest"], "generator output drifted from the fixture"
    return n, u, v, w


def problem_of(fx: dict, mode: str = "reduced", **kw) -> Problem:
    n, u, v, w = graph_of(fx)
    return Problem.from_edges(n, u, v, w, reduce_degree_two=(mode == "reduced"), **kw)


def words_of(bits_words, W):
    """Fixture word list -> uint64 array of W words."""
    a = np.zeros(W, dtype=np.uint64)
    a[: len(bits_words)] = np.array([int(x) for x in bits_words], dtype=np.uint64)
    return a


def dot(a: np.ndarray, b: np.ndarray) -> int:
    x = np.bitwise_and(a, b)
    return int(sum(bin(int(t)).count("1") for t in x) & 1)
