"""Helpers shared by the parity tests: render a scene on the GPU and with the
CPU oracle and compare bit for bit."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def assert_same(gpu_frame, gpu_depth, gpu_winner, ref, mask=None, label=""):
    """Bit-exact comparison of frame (top-down), depth and winner (bottom-up).
    ``mask`` (H, W) bool, bottom-up, restricts the comparison."""
    H, W = gpu_depth.shape
    if mask is None:
        mask = np.ones((H, W), bool)
    gf = gpu_frame[::-1][mask]
    rf = ref.frame[::-1][mask]
    gd = gpu_depth[mask].view(np.uint32)
    rd = ref.depth[mask].view(np.uint32)
    gw = gpu_winner[mask]
    rw = ref.winner[mask]
    bad_w = np.count_nonzero(gw != rw)
    bad_d = np.count_nonzero(gd != rd)
    bad_f = np.count_nonzero(np.any(gf != rf, axis=-1))
    if bad_w or bad_d or bad_f:
        idx = np.argwhere(mask)[np.flatnonzero((gw != rw) | (gd != rd) | np.any(gf != rf, axis=-1))[:5]]
        detail = []
        for (y, x) in idx:
            detail.append(f"(x={x},y={y}) gpu id={gpu_winner[y, x]} d={gpu_depth[y, x]!r} bgra={tuple(gpu_frame[H-1-y, x])}"
                          f" | ref id={ref.winner[y, x]} d={ref.depth[y, x]!r} bgra={tuple(ref.frame[H-1-y, x])}")
        raise AssertionError(f"{label}: {bad_w} winner, {bad_d} depth, {bad_f} colour mismatches of "
                             f"{int(mask.sum())} pixels\n" + "\n".join(detail))


def max_lsb(gpu_frame, ref_frame):
    return int(np.max(np.abs(gpu_frame.astype(np.int16) - ref_frame.astype(np.int16))))


def grid_mask(H, W, grid):
    sx, sy, ox, oy = grid
    m = np.zeros((H, W), bool)
    m[oy::sy, ox::sx] = True
    return m


def oracle(scene, **kw):
    r = O.render(scene, **kw)
    assert r.rc == 0, f"oracle failed: {O.ERRORS.get(r.rc, r.rc)}"
    return r
