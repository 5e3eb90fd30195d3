"""Run statistics in the reference's stats CSV format (include/stats.h:256-338).

The reference appends one row per run to the file named on its command line, writing the header first when the
file does not exist yet.  Two layouts exist: the CPU drivers' 16 columns and the GPU driver's 22 columns
(``is_gpu_timings``).  Column names and their order, and the printf widths of the values (``%5d`` integers, ``%15lf``
seconds and megabytes), are kept.  The reference's header and GPU row also contain the indentation that its string
literals pick up from source line continuations; that whitespace is not reproduced, so readers that strip spaces
around each field (as :func:`read_stats_csv` does) see the same table.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

CPU_COLUMNS = ("Total_Nodes", "Total_Edges", "New_Edges", "Node_removed", "Fvs_size", "Initial_Cycles",
               "Final_cycles", "Total_Weight", "Construction_trees(s)", "Collect_cycles(s)", "Inspection_time(s)",
               "Precompute_SP(s)", "Independence_test(s)", "Preprocessing Time(s)", "Main_loop(s)", "Total_time(s)")

GPU_COLUMNS = ("Total_Nodes", "Total_Edges", "New_Edges", "Nodes_removed", "Fvs_size", "Initial_Cycles",
               "Final_cycles", "Total_Weight", "nchunks", "nstreams", "total_memory_usage(mb)",
               "static_memory_usage(mb)", "variable_memory_usage(mb)", "load_entire_memory", "Construction_trees(s)",
               "Collect_cycles(s)", "Inspection_time(s)", "Gpu_timings", "Hybrid_Timings(s)", "Preprocessing Time(s)",
               "Main_loop(s)", "Total_time(s)")


def int32(x: int) -> int:
    """The reference accumulates the weight in an int (include/stats.h:19)."""
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x >= 1 << 31 else x


@dataclass
class RunStats:
    """The fields of one row.  Edge counts are directed (the reference counts CSR entries, 2 per edge)."""

    total_nodes: int
    total_edges: int
    new_edges: int
    nodes_removed: int
    fvs_size: int
    initial_cycles: int
    final_cycles: int
    total_weight: int                    # exact; written as the reference's int32
    construction_trees_s: float = 0.0    # host: graph, FVS, shortest-path trees
    collect_cycles_s: float = 0.0        # host: candidate filter and ranking
    inspection_s: float = 0.0            # the phase loop
    precompute_s: float = 0.0            # CPU layout: parity time; GPU layout: device time
    independence_s: float = 0.0          # CPU layout: support-vector update time
    nchunks: int = 1
    nstreams: int = 1
    total_memory_mb: float = 0.0
    static_memory_mb: float = 0.0
    variable_memory_mb: float = 0.0
    load_entire_memory: bool = True
    extra: dict = field(default_factory=dict)

    @property
    def preprocessing_s(self) -> float:
        return self.construction_trees_s + self.collect_cycles_s

    def row(self, gpu: bool) -> str:
        ints = [self.total_nodes, self.total_edges, self.new_edges, self.nodes_removed, self.fvs_size,
                self.initial_cycles, self.final_cycles, int32(self.total_weight)]
        if not gpu:
            secs = [self.construction_trees_s, self.collect_cycles_s, self.inspection_s, self.precompute_s,
                    self.independence_s, self.preprocessing_s, self.inspection_s,
                    self.inspection_s + self.preprocessing_s]
            return ",".join([f"{v:5d}" for v in ints] + [f"{v:15f}" for v in secs]) + "\n"
        ints += [self.nchunks, self.nstreams]
        mem = [self.total_memory_mb, self.static_memory_mb, self.variable_memory_mb]
        secs = [self.construction_trees_s, self.collect_cycles_s, self.inspection_s, self.precompute_s,
                self.independence_s, self.preprocessing_s, self.inspection_s, self.inspection_s + self.preprocessing_s]
        return ",".join([f"{v:5d}" for v in ints] + [f"{v:15f}" for v in mem] +
                        ["YES" if self.load_entire_memory else "NO"] + [f"{v:15f}" for v in secs]) + "\n"


def append_stats_csv(path: str, stats: RunStats, gpu: bool = True) -> None:
    """Append one row, writing the header first if the file does not exist (stats.h:256-338)."""
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write(",".join(GPU_COLUMNS if gpu else CPU_COLUMNS) + "\n")
        f.write(stats.row(gpu))


def read_stats_csv(path: str) -> list[dict]:
    """Rows of a stats CSV (ours or the reference's) as dicts keyed by the stripped column names."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    if not lines:
        return []
    names = [c.strip() for c in lines[0].split(",")]
    rows = []
    for ln in lines[1:]:
        vals = [v.strip() for v in ln.split(",")]
        row = {}
        for k, v in zip(names, vals):
            try:
                row[k] = int(v)
            except ValueError:
                try:
                    row[k] = float(v)
                except ValueError:
                    row[k] = v
        rows.append(row)
    return rows


def stats_of(problem, total_weight: int, **times) -> RunStats:
    """The row of a run on ``problem`` (a host.Problem): counts from the problem, the given weight and times.
    A single-cycle graph is reported as the reference does (mcb/mcb_cycle_deg_two_remove.cpp:103-110)."""
    i = problem.info
    if i.single_cycle:
        return RunStats(i.input_nodes, 2 * i.input_edges, 0, i.input_nodes - 1, 1, 1, 1, total_weight, **times)
    return RunStats(i.input_nodes, 2 * i.input_edges, i.directed_edges, i.nodes_removed, i.fvs_size,
                    i.num_candidates, i.num_non_tree, total_weight, **times)
