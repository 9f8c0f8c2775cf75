"""Launch-shape autotuner of the queue-fed advance kernel.

The kernel's results do not depend on its launch shape (the per-cell work and
operation order are fixed; only which lane takes which cell changes), so the
shape is chosen by timing alone: threads per block (the cells a block holds
at once), resident blocks per CU (0 = the occupancy limit of the kernel's
register use) and whether the first wave of block 0 takes the 64 cells that
needed the most attempts in the last step (running alone on its CU when the
block has 256 threads), and how many of the most expensive cells of each
step are advanced by whole waves, one cell per wave (wr_set_coop; the
lane-group width of the path is then 64 lanes for those cells and 1 for the
rest).  Each candidate times a short run of the caller's workload from the
same initial cells (after an untimed one that captures the HIP graph), and
the fastest is left set on the context.
"""
from __future__ import annotations

import itertools
import time

# coop_cells -1 is the context's default (one cooperative cell per 2,048 cells,
# at most 128).  The candidates' short runs cover the first outer steps, where
# the most expensive cells take few attempts; the cooperative waves pay off in
# the later steps, in which one cell takes tens of attempts, so the default is
# kept rather than tuned (64^3 x 1000 steps: 8-128 cooperative cells 2.71-2.81 s,
# none 4.39 s).
DEFAULT_SPACE = {"threads": (64, 128, 256), "blocks_per_cu": (0, 1), "solo_lanes": (0, 64),
                 "coop_cells": (-1,)}


def candidates(space=None):
    sp = dict(DEFAULT_SPACE, **(space or {}))
    for th, pc, so, co in itertools.product(sp["threads"], sp["blocks_per_cu"], sp["solo_lanes"],
                                            sp["coop_cells"]):
        if so <= th:
            yield th, pc, so, co


def autotune(ctx, cells, steps=100, perturb=True, space=None, repeats=1):
    """Time `steps` outer steps of `cells` for every candidate shape; set and
    return the fastest as (threads, blocks_per_cu, solo_lanes, coop_cells), with
    the table of (shape, seconds, grid blocks)."""
    table = []
    for shape in candidates(space):
        ctx.set_queue_launch(*shape[:3])
        ctx.set_coop(shape[3])
        ctx.upload(cells)
        ctx.run(perturb=perturb, max_steps=8)  # graph capture and occupancy query, untimed
        best = float("inf")
        for _ in range(repeats):
            ctx.upload(cells)
            t0 = time.perf_counter()
            ctx.run(perturb=perturb, max_steps=steps)
            best = min(best, time.perf_counter() - t0)
        table.append({"threads": shape[0], "blocks_per_cu": shape[1], "solo_lanes": shape[2],
                      "coop_cells": ctx.coop()[0], "seconds": best,
                      "grid_blocks": ctx.queue_launch()[3]})
    win = min(table, key=lambda r: r["seconds"])
    chosen = (win["threads"], win["blocks_per_cu"], win["solo_lanes"], win["coop_cells"])
    ctx.set_queue_launch(*chosen[:3])
    ctx.set_coop(chosen[3])
    return chosen, table
