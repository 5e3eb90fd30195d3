"""GPU tests of the graph-captured loops (mcb_solver_run_graph, mcb_batch_run_graph): the whole phase
loop recorded once as a HIP graph and replayed after each reset.  Every replay is checked against the
stream-launched run of the same problem (itself bit-exact against the CPU oracle, which is checked once
more here), with eager work on the device and host synchronisations between the replays."""
import numpy as np
import pytest
import torch

from minimumcyclebasis_amd import generators as G
from minimumcyclebasis_amd.components import split_components
from minimumcyclebasis_amd.device import Batch, Solver
from minimumcyclebasis_amd.host import Problem
from oracle import pyoracle

pytestmark = pytest.mark.gpu

REPLAYS = 4


def _eager_noise(i):
    """Unrelated device work between replays."""
    x = torch.arange(1 << 20, device="cuda", dtype=torch.int64) * (i + 3)
    torch.cuda.synchronize()
    return int(x.sum().item())


@pytest.mark.parametrize("case,block_phases", [("er", 0), ("er", 1), ("delaunay", 0), ("grid", 0)])
def test_solver_graph_replays(case, block_phases):
    gr = {"er": lambda: G.erdos_renyi_lbc(1500, 5000, 1000, seed=11),
          "delaunay": lambda: G.delaunay(2500, seed=12),
          "grid": lambda: G.grid(30, 25, 50, seed=13)}[case]()
    p = Problem.from_edges(*gr, keep_vertex_trees=True, prune_trees=True)
    print(f"case solver graph {case} block_phases={block_phases} N={p.info.num_non_tree}", flush=True)
    ref = Solver(p, store_cycles=True, block_phases=block_phases)
    ref.run()
    ref.sync()
    want_sel, want_c = ref.selected(), ref.cycle_vectors()
    want_work = ref.work_counters()
    ref.close()
    assert np.array_equal(want_sel, pyoracle.run(p)["selected"])
    s = Solver(p, store_cycles=True, block_phases=block_phases)
    for i in range(REPLAYS):
        if i:
            s.reset()
        s.run_graph()
        assert s.phase == p.info.num_non_tree
        s.sync()
        assert np.array_equal(s.selected(), want_sel), f"replay {i}"
        assert np.array_equal(s.cycle_vectors(), want_c), f"replay {i}"
        if block_phases != 1:  # the one-kernel-per-phase search stops early at a timing-dependent point
            assert s.work_counters()[:2] == want_work[:2], f"replay {i}"
        _eager_noise(i)
    # the stream-launched loop still runs on a solver that holds a graph
    s.reset()
    s.run()
    s.sync()
    assert np.array_equal(s.selected(), want_sel)
    s.close()


def test_solver_graph_needs_phase_zero_and_no_timing():
    p = Problem.from_edges(*G.grid(10, 9, 50, seed=1), keep_vertex_trees=True, prune_trees=True)
    s = Solver(p)
    s.run(1)
    with pytest.raises(RuntimeError):
        s.run_graph()
    s.close()
    t = Solver(p, timing=True)
    with pytest.raises(RuntimeError):
        t.run_graph()
    t.close()


def test_batch_graph_replays():
    n, u, v, w, _ = G.batched_components(k=40, nmin=30, nmax=400, seed=21)
    probs = []
    for comp in split_components(n, u, v, w):
        p = Problem.from_edges(comp.n, comp.u, comp.v, comp.w, prune_trees=True, keep_vertex_trees=True)
        if not p.info.single_cycle:
            probs.append(p)
    print(f"case batch graph problems={len(probs)}", flush=True)
    ref = Batch(probs, store_cycles=True)
    ref.run()
    ref.sync()
    want = [(ref.selected(j), ref.cycle_vectors(j)) for j in range(len(probs))]
    ref.close()
    for j in (0, len(probs) // 2, len(probs) - 1):
        assert np.array_equal(want[j][0], pyoracle.run(probs[j])["selected"]), f"problem {j}"
    b = Batch(probs, store_cycles=True)
    for i in range(REPLAYS):
        if i:
            b.reset()
        b.run_graph()
        assert b.step == b.steps
        b.sync()
        for j in range(len(probs)):
            assert np.array_equal(b.selected(j), want[j][0]), f"replay {i} problem {j}"
            assert np.array_equal(b.cycle_vectors(j), want[j][1]), f"replay {i} problem {j}"
        _eager_noise(i)
    b.close()
