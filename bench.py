#!/usr/bin/env python3
"""MCB support-vector phase loop benchmark on one MI355X.

A step is one complete de Pina phase loop (all N phases: tree-path parity + minimum-rank odd
candidate search, cycle vector, GF(2) support-vector update) over one synthetic graph whose
trees and candidates are already resident in HBM.  The default workload is the 100 000-point
planar Delaunay graph, timed in two forms side by side: plain (the reference's FVS order) and
learned (the FVS order of the learned vertex scorer); each form's host preprocessing and upload
are timed once and reported with its end-to-end time.  The roots' shortest-path trees are searched on
the GPU by default (``--trees host``: by the host library; the plain form also times that build and
checks the arrays identical).  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload delaunay100k] [--fvs both|plain|learned]
                    [--trees device|host] [--no-prune] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from minimumcyclebasis_amd import generators as G  # noqa: E402

METRIC = "MCB wall-time & support-phase HBM GB/s; batched+learned-FVS vs plain, exact"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # the planar target configuration (fits one GPU): the default
    "delaunay100k": ("Delaunay triangulation of 100000 uniform points, weight round(1000*length), seed 1",
                     lambda: G.delaunay(100000, seed=1)),
    "er20000": ("Erdos-Renyi G(n=20000, m=60000), largest biconnected component, weights U{1..1000}, seed 1",
                lambda: G.erdos_renyi_lbc(20000, 60000, 1000, seed=1)),
    "grid64": ("64x64 grid, weights U{1..100}, seed 1", lambda: G.grid(64, 64, 100, seed=1)),
    "chunglu50000": ("Chung-Lu n=50000 gamma 2.3 avg degree 8, LBC, weights U{1..10000}, seed 1",
                     lambda: G.chung_lu_lbc(50000, 2.3, 8.0, 10000, seed=1)),
}

# graphs of many biconnected components: solved component by component, batched (one arena) and plain
# (one solver per component, one after the other) timed side by side
COMPONENT_WORKLOADS = {
    "components2000": ("2000 random 2-connected components, vertex count log-uniform in [50, 5000], average "
                       "degree U[4, 8], weights U{1..1000}, joined by 1999 bridges, seed 1",
                       lambda: G.batched_components(k=2000, nmin=50, nmax=5000, seed=1)[:4]),
    "components200": ("200 random 2-connected components, vertex count log-uniform in [50, 2000], average "
                      "degree U[4, 8], weights U{1..1000}, joined by 199 bridges, seed 1",
                      lambda: G.batched_components(k=200, nmin=50, nmax=2000, seed=1)[:4]),
}


def host_cores() -> int:
    """Host threads for the CPU legs: the affinity mask, capped by OMP_NUM_THREADS when set."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(n, u, v, w, sel: np.ndarray, S: np.ndarray, threads: int, prune: bool, phases: int = 24,
                 roots: int = 4096) -> dict:
    """The CPU oracle (the reference's CPU phase loop, mcb/mcb_cycle.cpp:237-301, restated in C) timed on
    ``phases`` phases spread evenly over the whole loop, each started from the GPU's own S_e (``S``, every
    support vector as used in its phase; ``sel`` the GPU's selections, erased from the list as the reference
    erases them).  Per sampled phase: the per-vertex parity BFS of ``roots`` roots spread over the FVS (scaled by
    F / roots: every root's BFS covers its whole tree every phase), the list scan up to the first odd candidate,
    and the cycle vector with the products of every later vector over all W words.  The full-loop time is the
    mean scaled parity time times N plus the scan and product times integrated over the phases (trapezoid
    rule).  The XORs of the update are inside the timed call but run on the GPU's final later vectors, whose
    products with C_e are all even, so they cost nothing here (the reference XORs where a product is odd)."""
    from minimumcyclebasis_amd.host import Problem
    from oracle import pyoracle

    t_build = time.perf_counter()
    p = Problem.from_edges(n, u, v, w, prune_trees=prune, keep_vertex_trees=True, num_threads=threads)
    t_build = time.perf_counter() - t_build
    info = p.info
    N, F, nv, K = info.num_non_tree, info.fvs_size, info.nodes, info.num_candidates
    if S.shape != (N, info.words) or sel.size != N:
        raise ValueError("the GPU's support vectors and selections do not match the problem")
    to = p.array("tree_offset")
    flat = (np.repeat(np.arange(F, dtype=np.int64), np.diff(to)) * nv + p.array("pos_vertex").astype(np.int64))
    P = np.zeros(F * nv, dtype=np.uint8)
    sample = np.unique(np.linspace(0, N - 1, min(phases, N)).astype(np.int64))
    rs = np.unique(np.linspace(0, F - 1, min(roots, F)).astype(np.int32))
    t_par, t_scan, t_upd = [], [], []
    for e in sample:
        # the per-vertex parity of every root at the candidates' endpoints, untimed (the sampled roots' BFS,
        # timed, then covers their whole trees)
        P[flat] = pyoracle.position_parity(p, S[e], num_threads=threads)
        later = S[e + 1:]
        k, a, b, c = pyoracle.phase_timed(p, int(e), S[e], sel[:e], P, later, rs, num_threads=threads)
        if k != sel[e]:
            raise RuntimeError(f"CPU oracle chose rank {k} in phase {e}, the GPU {sel[e]}")
        t_par.append(a * F / rs.size)
        t_scan.append(b)
        t_upd.append(c)
    rest = np.asarray(t_scan) + np.asarray(t_upd)
    total = float(np.mean(t_par)) * N
    if sample.size > 1:  # sum over e = 0..N-1 of the piecewise-linear interpolation of the samples
        total += float(np.sum((rest[1:] + rest[:-1]) * np.diff(sample)) / 2.0 + (rest[0] + rest[-1]) / 2.0)
    else:
        total += float(rest[0]) * N
    spent = float(np.sum(t_par) * rs.size / F + np.sum(rest))
    p.close()
    return {"value": round(total, 1), "unit": "s per full phase loop (extrapolated)", "cores": threads,
            "kind": "port",
            "sample": f"{sample.size} phases evenly spread over all {N} (each started from the GPU's S_e, chosen "
                      f"rank checked equal), {spent:.2f} s timed: parity BFS of {rs.size} of {F} roots scaled by "
                      f"{F / rs.size:.1f}, list scan and products with all later vectors timed whole; per phase "
                      f"parity {np.mean(t_par):.3f} s, scan+products {rest.min():.3f}..{rest.max():.3f} s; "
                      f"vertex trees built in {t_build:.1f} s, untimed"}


def measured_traffic(kernel: str, workload: str | None, key: str = "traffic_bytes_per_launch:"):
    """HBM bytes per launch (or per step, by key) from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this kernel and workload (counter values scaled by the copy-kernel calibration), or None."""
    name = kernel if workload is None else f"{kernel}_{workload}"
    path = os.path.join(ROOT, "profiles", f"{name}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="delaunay100k", choices=sorted(WORKLOADS) + sorted(COMPONENT_WORKLOADS))
    ap.add_argument("--no-plain", action="store_true",
                    help="component workloads: skip the plain one-solver-per-component timing")
    ap.add_argument("--tree-waves", type=int, default=0,
                    help="tree pass waves per workgroup (4, 8, 16; 0 timed at create)")
    ap.add_argument("--tree-cap", type=int, default=0,
                    help="largest tree kept in LDS by the tree pass (0 timed at create, -1 as large as fits)")
    ap.add_argument("--merge-group", type=int, default=0,
                    help="first-level bases merged per wave of the greedy basis (0 automatic, 2..64)")
    ap.add_argument("--fvs", choices=["both", "plain", "learned"], default="both",
                    help="single graphs: FVS order(s) timed.  both: the plain (reference order) and the learned "
                         "form side by side, the learned one the headline; their weights must agree")
    ap.add_argument("--graph", action="store_true",
                    help="also time the loop with no HIP events: stream-launched, and replayed as one HIP graph")
    ap.add_argument("--no-prune", action="store_true")
    ap.add_argument("--trees", choices=["device", "host"], default="device",
                    help="where the roots' shortest-path trees are searched (identical results)")
    ap.add_argument("--no-host-trees", action="store_true",
                    help="skip the comparison build with host-searched trees (plain form)")
    ap.add_argument("--breakdown", action="store_true",
                    help="time every kernel (HIP events between all launches), not only the dominant one")
    ap.add_argument("--no-timing", action="store_true",
                    help="no HIP events at all (counter passes: the kernel times and rooflines come out empty)")
    ap.add_argument("--timing-every", type=int, default=16, metavar="K",
                    help="single graphs: HIP events around the timed kernels of every K-th block only (an event "
                         "pair holds the stream a few microseconds; 1 = every block)")
    ap.add_argument("--block-phases", type=int, default=0,
                    help="phases per pass over the trees (0 automatic, 1 one kernel per phase)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="component workloads: CPU oracle seconds")
    ap.add_argument("--cpu-phases", type=int, default=24, help="single graphs: phases the CPU oracle times")
    ap.add_argument("--phases", type=int, default=0,
                    help="profiling aid: run only the first PHASES phases per step (0 = the full loop)")
    ap.add_argument("--progress-phases", type=int, default=0, metavar="K",
                    help="profiling aid: run the loop K phases at a time, waiting and printing a line after each")
    ap.add_argument("--copy-kernel", type=int, default=0, metavar="BYTES",
                    help="calibration aid: run only the device copy kernel over BYTES bytes and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the phase loop of one graph runs on one device "
              "(multi-GPU sharding by root is a design note, not part of this build)", file=sys.stderr)
        return 2

    from minimumcyclebasis_amd.device import copy_bandwidth

    if args.copy_kernel:
        gbs = copy_bandwidth(args.copy_kernel, iters=max(1, args.steps))
        print(json.dumps({"copy_kernel_bytes": args.copy_kernel, "iters": max(1, args.steps), "gbps": round(gbs, 1)}),
              flush=True)
        return 0

    if args.workload in COMPONENT_WORKLOADS:
        return run_components(args)
    return run_single(args)


def time_leg(form: str, graph, args, threads: int) -> dict:
    """One form of the single-graph workload: host preprocessing (with the learned FVS order for
    form "learned"), upload (solver creation), then `warmup` untimed and `steps` timed phase loops."""
    from minimumcyclebasis_amd.device import KERNELS, Solver
    from minimumcyclebasis_amd.host import Problem

    n, u, v, w = graph
    t0 = time.perf_counter()
    fvs_info = None
    order = None
    if form == "learned":
        from minimumcyclebasis_amd.learned_fvs import learned_order

        ch = learned_order(n, u, v, w)
        order = ch.order
        fvs_info = {"fvs_reference": ch.fvs_reference, "fvs_learned": ch.fvs_learned, "learned_used": ch.used,
                    "learned_order_s": round(time.perf_counter() - t0, 3)}
    trees = "host" if args.no_prune else args.trees
    split = {}
    if trees == "device":
        from minimumcyclebasis_amd.device import problem_from_edges

        prob = problem_from_edges(n, u, v, w, trees="device", num_threads=threads, fvs_order=order, timings=split)
    else:
        prob = Problem.from_edges(n, u, v, w, prune_trees=not args.no_prune, num_threads=threads, fvs_order=order)
    host_s = time.perf_counter() - t0
    info = prob.info
    print(f"[bench] {args.workload} {form}: n'={info.nodes} m'={info.directed_edges // 2} N={info.num_non_tree} "
          f"F={info.fvs_size} K={info.num_candidates} positions={info.positions} preprocessing={host_s:.2f}s "
          f"(trees on the {trees}) {split}", file=sys.stderr, flush=True)
    host_trees_s = None
    if trees == "device" and form == "plain" and not args.no_host_trees:
        # the same problem with the trees searched by the host library, for comparison (arrays identical)
        t1 = time.perf_counter()
        ref = Problem.from_edges(n, u, v, w, prune_trees=True, num_threads=threads, fvs_order=order)
        host_trees_s = time.perf_counter() - t1
        same = all(np.array_equal(ref.array(a), prob.array(a)) for a in
                   ("cand_edge", "cand_root", "cand_weight", "tree_offset", "parent_pos", "pos_coord"))
        ref.close()
        if not same:
            raise RuntimeError("device and host shortest-path trees differ")
        print(f"[bench] {form}: host-searched trees {host_trees_s:.2f}s, identical arrays", file=sys.stderr,
              flush=True)
    t0 = time.perf_counter()
    timing = 0 if args.no_timing else 1 if args.breakdown else 2
    solver = Solver(prob, timing=timing, block_phases=args.block_phases, waves_per_block=args.tree_waves,
                    tree_cap=args.tree_cap, merge_group=args.merge_group, timing_every=args.timing_every)
    upload_s = time.perf_counter() - t0
    phases = args.phases if 0 < args.phases < info.num_non_tree else None

    def run_loop():
        if args.progress_phases <= 0:
            solver.run(phases)
            solver.sync()
            return
        end = phases if phases is not None else info.num_non_tree
        while solver.phase < end:
            solver.run(min(args.progress_phases, end - solver.phase))
            solver.sync()
            print(f"[bench] phase {solver.phase} of {end}", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        solver.reset()
        run_loop()
        print(f"[bench] {form} warmup {i} done", file=sys.stderr, flush=True)
    kt = {k: [0.0, 0] for k in KERNELS}
    pos = cand = upd = 0
    t_total = 0.0
    sel = None
    for i in range(args.steps):
        solver.reset()
        t0 = time.perf_counter()
        run_loop()
        dt = time.perf_counter() - t0
        t_total += dt
        for k in KERNELS:
            ms, cnt = solver.kernel_time(k)
            kt[k][0] += ms
            kt[k][1] += cnt
        p_, c_, u_ = solver.work_counters()
        pos += p_
        cand += c_
        upd += u_
        sel = solver.selected()
        print(f"[bench] {form} step {i}: {dt:.3f} s", file=sys.stderr, flush=True)
    loop_s = t_total / args.steps
    pre = {"trees": trees}
    if split:
        pre.update({k: round(v, 3) for k, v in split.items() if k != "search"})
        pre["search_device_ms"] = round(split["search"]["device_ms"], 1)
        pre["search_roots_per_stage"] = split["search"]["stage_roots"]
        pre["search_roots_left_to_host"] = split["search"]["host_roots"]
    if host_trees_s is not None:
        pre["host_trees_preprocess_s"] = round(host_trees_s, 3)
    return {"form": form, "prob": prob, "solver": solver, "host_s": host_s, "upload_s": upload_s, "loop_s": loop_s,
            "kt": kt, "pos": pos, "cand": cand, "upd": upd, "sel": sel, "shape": solver.launch_shape(),
            "grids": solver.update_grids(),
            "fvs": fvs_info, "total_weight": int(prob.array("cand_weight")[sel].astype(np.int64).sum()),
            "pre": pre}


def leg_summary(leg: dict) -> dict:
    info = leg["prob"].info
    out = {"fvs_size": info.fvs_size, "candidates": info.num_candidates, "tree_positions": info.positions,
           "host_preprocess_s": round(leg["host_s"], 3), "upload_s": round(leg["upload_s"], 3),
           "loop_s": round(leg["loop_s"], 6),
           "end_to_end_s": round(leg["host_s"] + leg["upload_s"] + leg["loop_s"], 3),
           "total_weight": leg["total_weight"], "preprocessing": leg["pre"]}
    if leg["fvs"]:
        out.update(leg["fvs"])
    return out


def run_single(args) -> int:
    from minimumcyclebasis_amd.device import Solver, device_info

    desc, gen = WORKLOADS[args.workload]
    graph = gen()
    threads = host_cores()
    forms = ["plain", "learned"] if args.fvs == "both" else [args.fvs]
    legs = {f: time_leg(f, graph, args, threads) for f in forms}
    head = legs[forms[-1]]
    dev = device_info()
    prob, info, steps = head["prob"], head["prob"].info, args.steps
    kt, pos, cand, upd = head["kt"], head["pos"], head["cand"], head["upd"]

    # Algorithmic bytes of one launch of the dominant kernel.  One kernel per phase: 8 per tree
    # position streamed (u32 parent position + i32 coordinate), 16 per candidate read (u_pos, v_pos,
    # rank, coordinate) and the update's support-word reads and read-modify-writes as the kernel
    # counts them.  Blocked: the tree pass streams every position (8 B) and every candidate (16 B);
    # its few independent candidate words are written sparsely and the update is a kernel of its
    # own, reported apart.
    # The work counters cover every launch; the events only the sampled blocks (every
    # --timing-every-th), whose average duration stands for all of them.
    shape = head["shape"]
    dominant = "k_phase" if shape["block_phases"] == 1 else "k_block_q"
    k_ms, k_n = kt["parity_search"]
    avg_ms = k_ms / max(k_n, 1)
    blocked = shape["block_phases"] > 1
    run_phases = args.phases if 0 < args.phases < info.num_non_tree else info.num_non_tree
    per_step = -(-run_phases // shape["block_phases"])  # launches of the dominant kernel per loop
    launches = per_step * steps
    cand_bytes = 16.0
    dom_bytes = 8.0 * pos + cand_bytes * cand + (0 if blocked else upd)
    bytes_per_launch = dom_bytes / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    loop_s = head["loop_s"]
    out = {
        "metric": METRIC,
        "value": round(loop_s, 6),
        "unit": "s per MCB phase loop",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * loop_s, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "graph": desc,
            "nodes_reduced": info.nodes,
            "edges_reduced": info.directed_edges // 2,
            "num_non_tree": info.num_non_tree,
            "fvs_size": info.fvs_size,
            "candidates": info.num_candidates,
            "tree_positions": info.positions,
            "pruned_trees": not args.no_prune,
            "form": "learned-fvs" if head["form"] == "learned" else "plain",
            "device": dev["name"],
        },
        "total_weight": head["total_weight"],
        "host_preprocess_s": round(head["host_s"], 3),
        "preprocessing": head["pre"],
        "upload_s": round(head["upload_s"], 3),
        "end_to_end_s": round(head["host_s"] + head["upload_s"] + loop_s, 3),
        "block_phases": shape["block_phases"],
        "tree_pass_waves": shape["waves_per_block"],
        "tree_pass_lds": shape["lds_bytes"],
        "update_grids": [head["grids"]["dot_workgroups"], head["grids"]["xor_workgroups"]],
        "launches_per_step": per_step,
        "timed_launches": k_n,
        "kernel_ms_per_step": {k: round(v[0] / v[1] * per_step, 3) for k, v in kt.items() if v[1]},
        "update_bytes_per_step": upd // max(steps, 1),
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dominant, args.workload) if not args.no_prune else None,
            "bytes_per_launch": round(bytes_per_launch, 1),
            "bytes_split": {"tree_positions": round(8.0 * pos / launches, 1),
                            "candidates": round(cand_bytes * cand / launches, 1),
                            "support_update": 0.0 if blocked else round(upd / launches, 1)},
            "avg_launch_us": round(avg_ms * 1000.0, 3),
        },
        "roofline_update": update_roofline(kt, upd, steps, per_step, "k_block_dot+k_block_xor",
                                           f"k_block_update_{args.workload}") if blocked and not args.no_prune else None,
    }
    if len(legs) == 2:
        pl, le = legs["plain"], legs["learned"]
        out["forms"] = {"plain": leg_summary(pl), "learned": leg_summary(le)}
        e2e = {f: out["forms"][f]["end_to_end_s"] for f in ("plain", "learned")}
        out["speedup_vs_plain"] = {"loop": round(pl["loop_s"] / le["loop_s"], 3),
                                   "end_to_end": round(e2e["plain"] / e2e["learned"], 3)}
        out["weights_equal"] = pl["total_weight"] == le["total_weight"]
        if not out["weights_equal"]:
            print("[bench] learned and plain FVS give different MCB weights", file=sys.stderr, flush=True)
            return 1
    elif head["fvs"]:
        out["learned_fvs"] = head["fvs"]
    if args.graph:
        sel = head["sel"]
        out["untimed"] = untimed_legs(
            lambda: Solver(prob, block_phases=args.block_phases, waves_per_block=args.tree_waves,
                           tree_cap=args.tree_cap),
            args.steps, lambda o: np.array_equal(o.selected(), sel))
    base = legs.get("plain", head)
    S = base["solver"].support_vectors() if not args.no_cpu_baseline else None
    for leg in legs.values():
        leg["solver"].close()
    if not args.no_cpu_baseline:
        # the reference's FVS order: the CPU oracle restates the reference's loop on the reference's root set
        if base["form"] != "plain":
            raise RuntimeError("the CPU baseline needs the plain form's loop")
        print("[bench] CPU baseline: sampled phases of the CPU oracle", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(*graph, base["sel"], S, threads, prune=not args.no_prune,
                                           phases=args.cpu_phases)
        del S
    for leg in legs.values():
        leg["prob"].close()
    print(json.dumps(out), flush=True)
    return 0


def update_roofline(kt, upd: int, steps: int, per_step: int, kernels: str, profile: str):
    """The support vector update (dot products and XORs of later vectors): the bytes the update
    kernels count as they run (each later vector's live prefix read once per block, and read and
    written once more where the block's combination is nonzero) over their event-timed time.  The
    events may cover a sample of the per_step blocks of a loop; their mean stands for every block."""
    ms, n = kt["update"]
    if n == 0 or ms <= 0:
        return None
    ms_step = ms / n * per_step
    gbs = upd / steps / (ms_step * 1e-3) / 1e9
    return {"kernels": kernels, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_step": upd // max(steps, 1),
            "ms_per_step": round(ms_step, 3), "launch_pairs_per_step": per_step, "timed_pairs": n,
            "traffic": measured_traffic(profile, None, "traffic_bytes_per_step:")}


def untimed_legs(make, steps: int, want) -> dict:
    """Loop times with no HIP events around any kernel: stream-launched (run) and as one HIP graph replay
    (run_graph).  `make()` builds a fresh solver or batch without timing; `want(obj)` checks its results
    against the timed run's.  Each leg runs once untimed first (the graph leg captures there)."""
    out = {}
    for leg in ("stream", "graph"):
        obj = make()
        go = obj.run_graph if leg == "graph" else obj.run
        go()
        obj.sync()
        t = 0.0
        for _ in range(steps):
            obj.reset()
            t0 = time.perf_counter()
            go()
            obj.sync()
            t += time.perf_counter() - t0
        if not want(obj):
            obj.close()
            raise RuntimeError(f"{leg} leg: results differ from the timed run")
        obj.close()
        out[f"{leg}_untimed_loop_s"] = round(t / steps, 6)
        print(f"[bench] {leg} untimed: {t / steps:.4f} s", file=sys.stderr, flush=True)
    return out


def cpu_baseline_components(comps, work, budget_s: float, threads: int) -> dict:
    """The CPU oracle (the reference's CPU phase loop restated, mcb/mcb_cycle.cpp:226-305) over whole
    components taken in a seeded random order until about budget_s is spent; the total is extrapolated by
    the oracle's work model (work[i]: phases x roots x vertices of component i, 0 for a single cycle)."""
    from minimumcyclebasis_amd.host import Problem
    from oracle import pyoracle

    rng = np.random.default_rng(5)
    order = rng.permutation(len(comps))
    work = np.asarray(work, dtype=np.float64)
    spent = done_work = 0.0
    done = 0
    median = float(np.median(work[work > 0])) if (work > 0).any() else 0.0
    for i in order:
        if work[i] == 0.0:
            continue
        # skip a component predicted to overrun the budget (before any rate is known: one above the median)
        if done_work == 0.0 and work[i] > median:
            continue
        if done_work > 0.0 and work[i] * spent / done_work > budget_s - spent:
            continue
        c = comps[i]
        p = Problem.from_edges(c.n, c.u, c.v, c.w, keep_vertex_trees=True, num_threads=threads)
        t0 = time.perf_counter()
        pyoracle.run(p, num_threads=threads)
        spent += time.perf_counter() - t0
        done_work += work[i]
        done += 1
        p.close()
        if spent >= budget_s:
            break
    total = spent * work.sum() / max(done_work, 1.0)
    return {"value": round(float(total), 3), "unit": "s per phase loop over all components (extrapolated)",
            "cores": threads, "kind": "port",
            "sample": f"{done} of {len(comps)} components (random order, whole loops) timed ({spent:.2f} s), "
                      f"scaled by phases x roots x vertices"}


def run_components(args) -> int:
    """Component workloads: split into biconnected components, preprocess each on the host, then time the
    batched loop (every component in one arena) and the plain one (one solver per component, in turn)."""
    from minimumcyclebasis_amd.components import build_problems, split_components
    from minimumcyclebasis_amd.device import KERNELS, Batch, Solver, device_info

    desc, gen = COMPONENT_WORKLOADS[args.workload]
    n, u, v, w = gen()
    threads = host_cores()
    t0 = time.perf_counter()
    comps = split_components(n, u, v, w)
    probs = build_problems(comps, threads=threads, prune_trees=not args.no_prune)
    host_s = time.perf_counter() - t0
    loop = [p for p in probs if not p.info.single_cycle]
    single = len(probs) - len(loop)
    Ns = [p.info.num_non_tree for p in loop]
    tot = {k: int(sum(getattr(p.info, k) for p in loop)) for k in ("num_non_tree", "num_candidates", "positions",
                                                                     "fvs_size")}
    print(f"[bench] {args.workload}: {len(comps)} components ({single} single cycles) N={tot['num_non_tree']} "
          f"K={tot['num_candidates']} positions={tot['positions']} F={tot['fvs_size']} host={host_s:.2f}s",
          file=sys.stderr, flush=True)
    dev = device_info()
    t0 = time.perf_counter()
    timing = 0 if args.no_timing else 1 if args.breakdown else 2
    bt = Batch(loop, timing=timing, waves_per_block=args.tree_waves, tree_cap=args.tree_cap,
               merge_group=args.merge_group)
    upload_s = time.perf_counter() - t0
    print(f"[bench] batch arena {bt.device_bytes / 2**30:.2f} GiB, {bt.steps} steps, upload {upload_s:.2f}s",
          file=sys.stderr, flush=True)
    for i in range(args.warmup):
        bt.reset()
        bt.run()
        bt.sync()
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    kt = {k: [0.0, 0] for k in KERNELS}
    pos = cand = upd = 0
    t_total = 0.0
    for i in range(args.steps):
        bt.reset()
        t0 = time.perf_counter()
        bt.run()
        bt.sync()
        dt = time.perf_counter() - t0
        t_total += dt
        for k in KERNELS:
            ms, cnt = bt.kernel_time(k)
            kt[k][0] += ms
            kt[k][1] += cnt
        p_, c_, u_ = bt.work_counters()
        pos += p_
        cand += c_
        upd += u_
        print(f"[bench] step {i}: {dt:.3f} s", file=sys.stderr, flush=True)
    batched_s = t_total / args.steps
    sels = [bt.selected(j) for j in range(len(loop))]
    total_weight = sum(int(p.array("cand_weight")[sel].astype(np.int64).sum()) for p, sel in zip(loop, sels))
    total_weight += sum(int(p.info.input_total_weight) for p in probs if p.info.single_cycle)
    shape = bt.launch_shape()
    bt.close()
    untimed = None
    if args.graph:
        untimed = untimed_legs(
            lambda: Batch(loop, waves_per_block=args.tree_waves, tree_cap=args.tree_cap), args.steps,
            lambda o: all(np.array_equal(o.selected(j), sels[j]) for j in range(len(loop))))

    plain_s = plain_upload_s = None
    if not args.no_plain:
        # the same components, one solver at a time, as the per-component driver runs them (creation timed apart)
        plain_s = plain_upload_s = 0.0
        for j, p in enumerate(loop):
            t0 = time.perf_counter()
            s_ = Solver(p)
            t1 = time.perf_counter()
            s_.run()
            s_.sync()
            plain_s += time.perf_counter() - t1
            plain_upload_s += t1 - t0
            if not np.array_equal(s_.selected(), sels[j]):
                print(f"[bench] component {j}: batched and plain selections differ", file=sys.stderr, flush=True)
                return 1
            s_.close()
        print(f"[bench] plain: {plain_s:.3f} s", file=sys.stderr, flush=True)

    k_ms, k_n = kt["parity_search"]
    avg_ms = k_ms / max(k_n, 1)
    bytes_per_launch = (8.0 * pos + 16.0 * cand) / max(k_n, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": round(batched_s, 6),
        "unit": "s per MCB phase loop over all components",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * batched_s, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "graph": desc,
            "components": len(comps),
            "single_cycle_components": single,
            "num_non_tree": tot["num_non_tree"],
            "largest_num_non_tree": max(Ns),
            "fvs_size": tot["fvs_size"],
            "candidates": tot["num_candidates"],
            "tree_positions": tot["positions"],
            "pruned_trees": not args.no_prune,
            "form": "batched",
            "device": dev["name"],
        },
        "total_weight": total_weight,
        "host_preprocess_s": round(host_s, 3),
        "upload_s": round(upload_s, 3),
        "steps_per_loop": shape["steps"],
        "tree_pass_waves": shape["waves_per_block"],
        "tree_pass_lds": shape["lds_bytes"],
        "plain_loop_s": None if plain_s is None else round(plain_s, 6),
        "speedup_vs_plain": None if plain_s is None else round(plain_s / batched_s, 2),
        "end_to_end_s": round(host_s + upload_s + batched_s, 3),
        "plain_upload_s": None if plain_upload_s is None else round(plain_upload_s, 3),
        "plain_end_to_end_s": None if plain_s is None else round(host_s + plain_upload_s + plain_s, 3),
        "speedup_vs_plain_end_to_end": None if plain_s is None else
        round((host_s + plain_upload_s + plain_s) / (host_s + upload_s + batched_s), 3),
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in kt.items() if v[1]},
        "update_bytes_per_step": upd // max(args.steps, 1),
        "roofline": {
            "kernel": "k_batch_q",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic("k_batch_q", args.workload) if not args.no_prune else None,
            "bytes_per_launch": round(bytes_per_launch, 1),
            "avg_launch_us": round(avg_ms * 1000.0, 3),
        },
        "roofline_update": update_roofline(kt, upd, args.steps, k_n // max(args.steps, 1), "k_batch_dot+k_batch_xor",
                                           f"k_batch_update_{args.workload}"),
    }
    if untimed is not None:
        out["untimed"] = untimed
    work = [0.0 if p.info.single_cycle else float(p.info.num_non_tree) * p.info.fvs_size * p.info.nodes
            for p in probs]
    for p in probs:
        p.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_components(comps, work, args.cpu_budget, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
