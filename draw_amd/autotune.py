"""Autotuner for the raster configuration: tile size and waves per tile.

Every supported (tile, waves) pair renders the caller's workload a few times;
the device time of each frame comes from HIP events on the context stream
(dh_get_stats ms_frame), and the pair with the lowest median wins.  Output does
not depend on the configuration (the per-pixel result is the same for every
tiling), only speed does.  Bin capacity needs no tuning: the library sizes
its bin buffers from the previous frame and regrows them on overflow.
"""
from __future__ import annotations

import statistics
from typing import Callable, Dict, List, Tuple

CANDIDATES: List[Tuple[int, int]] = [(8, 1), (16, 1), (16, 2), (16, 4), (32, 2), (32, 4), (32, 8), (32, 16)]


def autotune(renderer, render_once: Callable[[], object], candidates=CANDIDATES, frames: int = 3,
             warmup: int = 1) -> Tuple[Tuple[int, int], Dict[str, float]]:
    """Returns ((tile, waves), {"tile x waves": median ms}) and leaves the
    renderer set to the winner."""
    renderer.set_profiling(True)
    table: Dict[str, float] = {}
    best, best_ms = None, float("inf")
    for tile, waves in candidates:
        renderer.set_tiling(tile, waves)
        for _ in range(warmup):
            render_once()
        ms = []
        for _ in range(frames):
            render_once()
            ms.append(renderer.stats()["ms_frame"])
        med = statistics.median(ms)
        table[f"{tile}x{waves}"] = round(med, 4)
        if med < best_ms:
            best, best_ms = (tile, waves), med
    renderer.set_tiling(*best)
    return best, table
