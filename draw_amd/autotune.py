"""Autotuner for the raster configuration: raster kernel kind, tile size and
waves per tile.

Every candidate (kind, tile, waves) renders the caller's workload a few times;
the device time of each frame comes from HIP events on the context stream
(dh_get_stats ms_frame), and the pair with the lowest median wins.  Output does
not depend on the configuration (the per-pixel result is the same for every
tiling), only speed does.  Bin capacity needs no tuning: the library sizes
its bin buffers from the previous frame and regrows them on overflow.
"""
from __future__ import annotations

import statistics
from typing import Callable, Dict, List, Tuple

# kind 0: waves own 8x8 quadrants; kind 1: triangle-parallel rectangle sweeps
CANDIDATES: List[Tuple[int, int, int]] = [(0, 8, 1), (0, 16, 1), (0, 16, 2), (0, 16, 4), (0, 32, 2), (0, 32, 4),
                                          (0, 32, 8), (0, 32, 16), (1, 16, 4), (1, 16, 8), (1, 32, 4), (1, 32, 8)]


def autotune(renderer, render_once: Callable[[], object], candidates=CANDIDATES, frames: int = 3,
             warmup: int = 1) -> Tuple[Tuple[int, int, int], Dict[str, float]]:
    """Returns ((kind, tile, waves), {"k<kind>:<tile>x<waves>": median ms})
    and leaves the renderer set to the winner."""
    renderer.set_profiling(True)
    table: Dict[str, float] = {}
    best, best_ms = None, float("inf")
    for kind, tile, waves in candidates:
        renderer.set_raster(kind, tile, waves)
        for _ in range(warmup):
            render_once()
        ms = []
        for _ in range(frames):
            render_once()
            ms.append(renderer.stats()["ms_frame"])
        med = statistics.median(ms)
        table[f"k{kind}:{tile}x{waves}"] = round(med, 4)
        if med < best_ms:
            best, best_ms = (kind, tile, waves), med
    renderer.set_raster(*best)
    return best, table
