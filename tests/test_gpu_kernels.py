"""GPU parity tests: every kernel through the C ABI against the CPU oracle, bit-exact.

(a)  tree-path parity, native layout (Solver.parity) and the reference chunk layout (GpuStruct),
     against oracle_parity (include/work_per_thread.h:154-210 restated);
(a)+(b)+cycle+(c)  the whole phase loop: chosen rank per phase, every support vector S_e and every
     cycle vector C_e against oracle_run (mcb/mcb_cycle.cpp:226-305 restated)."""
import numpy as np
import pytest

from helpers import load_trace, problem_of, trace_names
from minimumcyclebasis_amd import _lib
from minimumcyclebasis_amd.device import GpuStruct, Solver
from oracle import pyoracle

pytestmark = pytest.mark.gpu

MODES = ("reduced", "plain")
PARITY_CASES = ["k4", "theta3", "grid5x5_ties", "delaunay40", "chunglu60", "grid16x16", "er150", "chunglu300",
                "biconn200", "delaunay150"]


def random_support(rng, N, W, density):
    S = np.zeros(W, dtype=np.uint64)
    bits = np.flatnonzero(rng.random(N) < density)
    for b in bits:
        S[b // 64] |= np.uint64(1) << np.uint64(b % 64)
    return S


@pytest.mark.parametrize("prune", [False, True])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", PARITY_CASES)
def test_parity_native_layout(name, mode, prune):
    fx = load_trace(name)
    p = problem_of(fx, mode, keep_vertex_trees=True, prune_trees=prune)
    print(f"case {name} {mode} prune={prune} F={p.info.fvs_size} P={p.info.positions}", flush=True)
    s = Solver(p)
    rng = np.random.default_rng(7)
    to, pv = p.array("tree_offset"), p.array("pos_vertex")
    root_of_pos = np.repeat(np.arange(p.info.fvs_size), np.diff(to))
    for density in (0.0, 0.05, 0.5, 1.0):
        S = random_support(rng, p.info.num_non_tree, p.info.words, density)
        got = s.parity(S)
        want = pyoracle.parity(p, S)
        assert np.array_equal(got, want[root_of_pos, pv]), f"density {density}"
    s.close()


@pytest.mark.parametrize("slots", [0, 1, 4])
@pytest.mark.parametrize("multiple_transfer", [False, True])
@pytest.mark.parametrize("chunk", [3, 720])
@pytest.mark.parametrize("name", ["grid4x4", "delaunay40", "chunglu60", "er150", "grid16x16"])
def test_parity_reference_layout(name, chunk, multiple_transfer, slots):
    """gpu_struct drop-in: compressed_trees chunk arrays in, per-position parity out.  slots > 0 caps the device
    chunk slots: with more chunks than slots the trees are streamed every phase (gpu/gpu_struct.cu:132-156,
    175-186), and a call without multiple_transfer is refused."""
    fx = load_trace(name)
    p = problem_of(fx, "reduced", keep_vertex_trees=True)
    n, F, N, W = p.info.nodes, p.info.fvs_size, p.info.num_non_tree, p.info.words
    print(f"case {name} chunk={chunk} mt={multiple_transfer} F={F} n={n}", flush=True)
    rows_c, cols_c, eo_c, ni_c, out_c = [], [], [], [], []
    nchunks = (F + chunk - 1) // chunk
    for i in range(nchunks):
        R = np.zeros(chunk * (n + 1), dtype=np.uint32)
        Cc = np.zeros(chunk * n, dtype=np.uint32)
        E = np.full(chunk * n, -7, dtype=np.int32)  # padded slots hold garbage, as pinned memory does
        NI = np.zeros(chunk * n, dtype=np.uint32)
        for s_ in range(chunk):
            r = i * chunk + s_
            if r >= F:
                break
            rr, cc, ee, nn = p.reference_layout(r)
            R[s_ * (n + 1):(s_ + 1) * (n + 1)] = rr
            Cc[s_ * n:(s_ + 1) * n] = cc
            E[s_ * n:(s_ + 1) * n] = ee
            NI[s_ * n:(s_ + 1) * n] = nn
        rows_c.append(R)
        cols_c.append(Cc)
        eo_c.append(E)
        ni_c.append(NI)
        out_c.append(np.full(chunk * n, 9, dtype=np.uint32))
    g = GpuStruct(p.info.directed_edges, N, W, n, F, chunk, nchunks, max_slots=slots)
    g.initialize_memory(p.array("nt_map"), rows_c, cols_c, eo_c)
    streamed = 0 < slots < nchunks
    if streamed and not multiple_transfer:
        with pytest.raises(_lib.MCBError):
            g.process_shortest_path(out_c, multiple_transfer=False)
        g.close()
        return
    rng = np.random.default_rng(3)
    for density in (0.1, 0.6, 0.3):
        S = random_support(rng, N, W, density)
        for o in out_c:
            o.fill(9)
        g.copy_support_vector(S)
        g.process_shortest_path(out_c, multiple_transfer=multiple_transfer)
        want = pyoracle.parity(p, S)
        for r in range(F):
            i, s_ = divmod(r, chunk)
            nodes_index = ni_c[i][s_ * n:(s_ + 1) * n]
            got = out_c[i][s_ * n:(s_ + 1) * n][nodes_index]
            assert np.array_equal(got, want[r].astype(np.uint32)), f"root {r}"
    g.close()


PHASE_CASES = trace_names()


# block_phases: 1 = one kernel per phase; 0 = automatic (64 phases per tree pass); 7 and 33 give
# partial words and blocks that end mid-word
BLOCKS = (1, 0, 7, 33)


@pytest.mark.parametrize("block", BLOCKS)
@pytest.mark.parametrize("prune", [False, True])
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", PHASE_CASES)
def test_phase_loop_matches_oracle(name, mode, prune, block):
    fx = load_trace(name)
    p = problem_of(fx, mode, keep_vertex_trees=True, prune_trees=prune)
    print(f"case {name} {mode} prune={prune} block={block} N={p.info.num_non_tree} K={p.info.num_candidates}",
          flush=True)
    ref = pyoracle.run(p, want_s=True, want_c=True)
    s = Solver(p, store_cycles=True, block_phases=block)
    assert s.launch_shape()["block_phases"] == (64 if block == 0 else block)
    s.run()
    s.sync()
    assert np.array_equal(s.selected(), ref["selected"])
    assert np.array_equal(s.support_vectors(), ref["S"])
    assert np.array_equal(s.cycle_vectors(), ref["C"])
    # the chosen candidates reproduce the reference's own chosen cycles and total
    r, a, b, w = p.candidate_table()
    sel = s.selected()
    chosen = np.stack([r[sel] + 1, a[sel] + 1, b[sel] + 1, w[sel]], axis=1).astype(np.int64).tolist()
    assert chosen == [ph["chosen"] for ph in fx[mode]["phases"]]
    s.close()


@pytest.mark.parametrize("block", BLOCKS)
def test_rerun_after_reset(block):
    fx = load_trace("grid16x16")
    p = problem_of(fx, "reduced", keep_vertex_trees=True)
    print(f"case rerun grid16x16 block={block}", flush=True)
    s = Solver(p, block_phases=block)
    s.run()
    s.sync()
    first = s.selected().copy()
    s.reset()
    s.run(50)
    s.run()
    s.sync()
    assert np.array_equal(s.selected(), first)


@pytest.mark.parametrize("block,pieces", [(0, (64, 10, 3, 100)), (7, (5, 7, 20, 1, 13)), (33, (40, 33, 2))])
@pytest.mark.parametrize("name", ["grid16x16", "er150", "chunglu300"])
def test_runs_in_pieces(name, block, pieces):
    """The loop enqueued in uneven pieces: a block cut short by the end of a run leaves the previous
    block's update of its next columns pending, and the next run applies it first (instead of on the
    fly); every selection, support vector and cycle vector still equals the oracle's."""
    fx = load_trace(name)
    p = problem_of(fx, "reduced", keep_vertex_trees=True, prune_trees=True)
    print(f"case pieces {name} block={block} pieces={pieces}", flush=True)
    ref = pyoracle.run(p, want_s=True, want_c=True)
    s = Solver(p, store_cycles=True, block_phases=block)
    for k in pieces:
        s.run(k)
    s.run()
    s.sync()
    assert np.array_equal(s.selected(), ref["selected"])
    assert np.array_equal(s.support_vectors(), ref["S"])
    assert np.array_equal(s.cycle_vectors(), ref["C"])
    s.close()


@pytest.mark.parametrize("cap", [-1, 64, 128])
@pytest.mark.parametrize("waves", [4, 8, 16])
@pytest.mark.parametrize("name", ["grid16x16", "er150", "delaunay150", "chunglu300", "biconn200"])
def test_tree_pass_widths(name, waves, cap):
    """Every width of the blocked tree pass (which trees go whole-workgroup and which one-wave changes
    with it), with the largest LDS-resident tree as large as fits or capped (larger trees in global
    scratch), gives the oracle's selections and cycle vectors."""
    fx = load_trace(name)
    p = problem_of(fx, "reduced", keep_vertex_trees=True, prune_trees=True)
    print(f"case {name} waves={waves} cap={cap}", flush=True)
    ref = pyoracle.run(p, want_c=True)
    s = Solver(p, store_cycles=True, waves_per_block=waves, tree_cap=cap)
    assert s.launch_shape()["waves_per_block"] == waves
    s.run()
    s.sync()
    assert np.array_equal(s.selected(), ref["selected"])
    assert np.array_equal(s.cycle_vectors(), ref["C"])
    s.close()


def _wheel(rim: int, seed: int):
    """A hub joined to every vertex of a rim cycle: every root's tree spans all vertices."""
    rng = np.random.default_rng(seed)
    r = np.arange(rim)
    u = np.concatenate([r, np.full(rim, rim)])
    v = np.concatenate([(r + 1) % rim, r])
    return rim + 1, u, v, rng.integers(1, 101, size=2 * rim)


@pytest.mark.parametrize("waves", [0, 4])
def test_tree_larger_than_lds_and_split_candidates(waves):
    """Trees of 25 001 positions (path words beyond the LDS, kept in global scratch) and roots with more
    than 8192 candidates (listed as several items) against the oracle."""
    from minimumcyclebasis_amd.host import Problem

    n, u, v, w = _wheel(25000, 3)
    p = Problem.from_edges(n, u, v, w, keep_vertex_trees=True, prune_trees=True)
    print(f"case wheel25000 waves={waves} F={p.info.fvs_size} max_pos={p.info.max_positions} "
          f"K={p.info.num_candidates}", flush=True)
    assert p.info.max_positions > 21000 and p.info.num_candidates > 8192
    ref = pyoracle.run(p, phases=700, want_c=True)
    s = Solver(p, store_cycles=True, waves_per_block=waves)
    assert s.launch_shape()["block_phases"] == 64
    s.run(700)
    s.sync()
    assert np.array_equal(s.selected(700), ref["selected"])
    assert np.array_equal(s.cycle_vectors(700), ref["C"])
    s.run()
    s.sync()
    r, a, b, wt = p.candidate_table()
    plain = Solver(p, block_phases=1)
    plain.run()
    plain.sync()
    assert int(wt[s.selected()].astype(np.int64).sum()) == int(wt[plain.selected()].astype(np.int64).sum())
    assert np.array_equal(s.selected(), plain.selected())
    s.close()
    plain.close()


_MERGE_CASES = {}


def _merge_case(case):
    """The problem and its oracle result, built once per session."""
    if case not in _MERGE_CASES:
        from minimumcyclebasis_amd import generators as G
        from minimumcyclebasis_amd.host import Problem

        gr = {"er1500": lambda: G.erdos_renyi_lbc(1500, 5000, 1000, seed=11),
              "delaunay2000": lambda: G.delaunay(2000, seed=4)}[case]()
        p = Problem.from_edges(*gr, keep_vertex_trees=True, prune_trees=True)
        _MERGE_CASES[case] = (p, pyoracle.run(p, want_c=True))
    return _MERGE_CASES[case]


@pytest.mark.parametrize("group", [0, 2, 3, 8, 64])
@pytest.mark.parametrize("case", ["er1500", "delaunay2000"])
def test_merge_group_sizes(case, group):
    """The greedy basis of a block merged in groups of 2, 3, 8 or 64 first-level bases (36 and 9 chunks of
    4096 ranks here, so up to six merge levels) gives the oracle's selections and cycle vectors."""
    p, ref = _merge_case(case)
    print(f"case {case} merge_group={group} K={p.info.num_candidates}", flush=True)
    s = Solver(p, store_cycles=True, merge_group=group)
    s.run()
    s.sync()
    assert np.array_equal(s.selected(), ref["selected"])
    assert np.array_equal(s.cycle_vectors(), ref["C"])
    s.close()


@pytest.mark.parametrize("n", [20000, 60000])
def test_first_level_rounds_agree_with_one_round(n):
    """By default a first-level basis of a large problem reads several 64-word rounds of the candidate
    bitmap (Delaunay 20k: about 500k candidates, three rounds; 60k: nine), so that one merge finishes the
    block; with a chosen merge group it reads one round and merges over two levels.  Both give the same
    selections and cycle vectors, and the sampled phases pass the oracle's per-phase checks."""
    from fullsize import check_loop, sample_phases
    from minimumcyclebasis_amd import generators as G
    from minimumcyclebasis_amd.host import Problem

    p = Problem.from_edges(*G.delaunay(n, seed=7), prune_trees=True)
    print(f"case delaunay{n} K={p.info.num_candidates}", flush=True)
    store = n <= 20000  # the cycle vectors of the larger case would take 2 GB per solver
    a = Solver(p, store_cycles=store)
    a.run()
    a.sync()
    b = Solver(p, store_cycles=store, merge_group=64)
    b.run()
    b.sync()
    assert np.array_equal(a.selected(), b.selected())
    if store:
        assert np.array_equal(a.cycle_vectors(), b.cycle_vectors())
    check_loop(p, a.selected(), a.support_vector_rows, sample_phases(p.info.num_non_tree, spread=10), seed=5,
               label=f"delaunay{n}")
    a.close()
    b.close()
    p.close()


def test_segmented_walk_resolves_agree():
    """Delaunay 10k with the tree pass's LDS trees capped at 768 positions: its 75 larger trees are narrow and
    walked in global scratch, a segment per wave of a 4-wave workgroup.  Resolving only each segment's last
    positions in order and joining the rest at the candidates' ends (the default), and resolving every segment
    in order, give the selections of the solver whose trees all stay in LDS; sampled phases pass the
    oracle's checks."""
    from fullsize import check_loop, sample_phases
    from minimumcyclebasis_amd import generators as G
    from minimumcyclebasis_amd.host import Problem

    p = Problem.from_edges(*G.delaunay(10000, seed=3), prune_trees=True)
    assert (np.diff(p.array("tree_offset")) > 768).sum() >= 50
    print(f"case delaunay10000 K={p.info.num_candidates}", flush=True)
    ref = Solver(p, waves_per_block=4, tree_cap=-1)
    ref.run()
    ref.sync()
    want = ref.selected()
    ref.close()
    for mode in (0, 1):
        print(f"case delaunay10000 cap 768 segment_resolve={mode}", flush=True)
        s = Solver(p, waves_per_block=4, tree_cap=768, segment_resolve=mode)
        s.run()
        s.sync()
        assert np.array_equal(s.selected(), want), f"segment_resolve={mode}"
        if mode == 0:
            check_loop(p, s.selected(), s.support_vector_rows, sample_phases(p.info.num_non_tree, spread=8),
                       seed=9, label="delaunay10000 segmented")
        s.close()
    with pytest.raises(RuntimeError):
        Solver(p, segment_resolve=2)
    p.close()


def test_update_grids_agree():
    """Every workgroup cap of the support vector update's two launches gives the same selections and
    support vectors: caps of one workgroup (every tile of a launch in turn), odd caps, the timed caps' range
    and caps above any launch's tile count, against the solver with the caps timed at create (one of the
    three timed values each); sampled phases pass the oracle's checks."""
    from fullsize import check_loop, sample_phases
    from minimumcyclebasis_amd import generators as G
    from minimumcyclebasis_amd.host import Problem

    small = problem_of(load_trace("grid16x16"), "reduced", keep_vertex_trees=True, prune_trees=True)
    p = Problem.from_edges(*G.delaunay(10000, seed=5), prune_trees=True)
    for prob, name, caps in ((small, "grid16x16", ((1, 1), (3, 5))),
                             (p, "delaunay10000", ((7, 3), (512, 4096), (65536, 65536)))):
        print(f"case {name} update grids timed", flush=True)
        ref = Solver(prob)
        g = ref.update_grids()
        assert g["dot_workgroups"] in (512, 1024, 2048) and g["xor_workgroups"] in (1024, 2048, 4096), g
        ref.run()
        ref.sync()
        want = ref.selected()
        rows = np.unique(np.linspace(0, prob.info.num_non_tree - 1, 40).astype(np.int32))
        want_s = ref.support_vector_rows(rows)
        ref.close()
        for d, x in caps:
            print(f"case {name} dot_workgroups={d} xor_workgroups={x}", flush=True)
            s = Solver(prob, dot_workgroups=d, xor_workgroups=x)
            assert s.update_grids() == {"dot_workgroups": d, "xor_workgroups": x}
            s.run()
            s.sync()
            assert np.array_equal(s.selected(), want), (name, d, x)
            assert np.array_equal(s.support_vector_rows(rows), want_s), (name, d, x)
            if (d, x) == (7, 3):
                check_loop(prob, s.selected(), s.support_vector_rows,
                           sample_phases(prob.info.num_non_tree, spread=8), seed=11, label=f"{name} grids {d},{x}")
            s.close()
    for bad in ((-1, 0), (0, 65537)):
        with pytest.raises(RuntimeError):
            Solver(small, dot_workgroups=bad[0], xor_workgroups=bad[1])
    small.close()
    p.close()


def test_merge_group_rejected():
    p = problem_of(load_trace("grid16x16"), "reduced", keep_vertex_trees=True, prune_trees=True)
    for bad in (1, 65, -2):
        with pytest.raises(RuntimeError):
            Solver(p, merge_group=bad)


@pytest.mark.parametrize("timing,every", [(2, 3), (1, 5), (2, 1)])
def test_sampled_timing(timing, every):
    """Events on every `every`-th block only: that many timed launches per kind, counted again from 0
    after a reset, and the selections of an untimed run."""
    p, ref = _merge_case("er1500")
    blocks = -(-p.info.num_non_tree // 64)
    print(f"case er1500 timing={timing} every={every} blocks={blocks}", flush=True)
    s = Solver(p, timing=timing, timing_every=every)
    kinds = ("parity_search", "update") if timing == 2 else ("parity_search", "select", "cycle", "update", "transpose")
    for _ in range(2):
        s.reset()
        s.run()
        s.sync()
        assert np.array_equal(s.selected(), ref["selected"])
        for k in kinds:
            ms, n = s.kernel_time(k)
            assert n == -(-blocks // every), k
            assert ms > 0, k
    s.close()
    with pytest.raises(RuntimeError):
        Solver(p, timing=2, timing_every=-1)
