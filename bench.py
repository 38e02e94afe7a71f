#!/usr/bin/env python3
"""Benchmark of the per-frame raster path on one MI355X.

Workload (default): the plain GPU path at 1920x1080 on the seeded 1M-triangle
soup (draw_amd.workloads.soup_scene(2)): one dh_render per frame, each frame
complete on the device before the next (one-frame-per-launch mode).  Inputs
(meshes, texture, canvas) are resident in HBM before the timed region.

Prints ONE JSON line: shaded Mpix/s (pixels with a depth-test winner per
second) as ``value``, plus triangles/s, p50/p99 frame latency, the frame time
with the canvas read back to the host, per-stage device times, the roofline of
the raster (the dominant kernel, which shades its own tiles in the default
raster kind), the 640x480 torus and 3840x2160 20M-triangle workloads, the
batched mode against the plain path on 256 views, the serving workload, and
the CPU oracle timed on whole frames of the same workloads on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, published


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def visible_triangles(winner: np.ndarray) -> tuple:
    """(covered pixels, distinct winning triangles) of one canvas."""
    w = winner.ravel()
    cov = w[w != 0xFFFFFFFF]
    if cov.size == 0:
        return 0, 0
    return int(cov.size), int(np.count_nonzero(np.bincount(cov)))


def shade_bytes(P: int, p_cov: int, t_vis: int) -> dict:
    """Algorithmic HBM bytes of one k_shade launch: per pixel a 4 B winner id
    read and a 4 B BGRA8 write; per covered pixel the 4 B diffuse texel (the
    ambient map is 1x1); per visible triangle 3 vertices x 14 f32 (normal,
    light, halfway, uv, depth).  ``record`` is what this implementation moves
    into the bench's canvases (no winner buffer): it also writes the f32 depth
    per pixel and reads a 96 B raster record plus a 144 B shading record per
    visible triangle."""
    return {"algorithmic": 8 * P + 4 * p_cov + 168 * t_vis, "record": 12 * P + 4 * p_cov + 240 * t_vis}


def raster_bytes(P: int, entries: int) -> dict:
    """Algorithmic HBM bytes of one k_raster launch: per pixel an f32 depth and
    a u32 id written; per bin entry a 4 B id and a 48 B setup record read (6
    snapped coordinates, 3 depths, 3 normalizers).  ``record``: this
    implementation writes only the 4 B winner per pixel and reads a 96 B record
    per entry."""
    return {"algorithmic": 8 * P + 52 * entries, "record": 4 * P + 100 * entries}


def raster_shade_bytes(P: int, entries: int, p_cov: int, t_vis: int) -> dict:
    """Algorithmic HBM bytes of one kind-0 k_raster launch, which shades its own
    tiles: per pixel a 4 B BGRA8 and an f32 depth written; per bin entry a 4 B
    id and a 48 B setup record read; per visible triangle 3 vertices x 14 f32
    and per covered pixel the 4 B diffuse texel.  ``record``: this
    implementation reads a 96 B raster record per entry and 224 B of records
    per visible triangle (the bench's canvases have no winner buffer; one that
    keeps winners adds 4 B per pixel)."""
    return {"algorithmic": 8 * P + 52 * entries + 168 * t_vis + 4 * p_cov,
            "record": 8 * P + 100 * entries + 224 * t_vis + 4 * p_cov}


TRAFFIC_PROFILE = "profiles/traffic_soup1m_1080p.md"
COUNTERS_PROFILE = "profiles/counters_soup1m_1080p.md"


def profile_rows(rel: str) -> dict:
    """The kernel rows of a committed rocprofv3 summary table: {kernel: cells}."""
    path = os.path.join(ROOT, rel)
    rows = {}
    if not os.path.exists(path):
        return rows
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 3 and cells[0].startswith("`"):
            rows[cells[0].strip("`").replace("dh::", "")] = cells[1:]
    return rows


def measured_traffic(kernel: str):
    """L2-to-memory bytes per launch of ``kernel`` on this workload, from the
    committed rocprofv3 --pmc summary: read requests by size
    (TCC_EA0_RDREQ_32B/64B/128B_sum x 32/64/128 B) plus write requests
    (TCC_EA0_WRREQ_64B_sum x 64 B, the rest x 32 B), per dispatch.  These
    counters read exactly 1 GiB each way on a 1 GiB device copy, so no scale
    factor is applied.  Returns (bytes, source) or (None, None)."""
    row = profile_rows(TRAFFIC_PROFILE).get(kernel)
    if not row:
        return None, None
    try:
        return int(float(row[-1])), TRAFFIC_PROFILE + " (TCC_EA0 read/write requests by size, per dispatch)"
    except ValueError:
        return None, None


def issue_roofline(kernel: str, ms: float):
    """VALU issue of ``kernel`` against the chip's: wave-level VALU
    instructions per launch (SQ_INSTS_VALU from the committed counter summary)
    over the launch time, against 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per
    wave64 VALU instruction = 614.4 G instructions/s."""
    row = profile_rows(COUNTERS_PROFILE).get(kernel)
    if not row or not ms:
        return None
    try:
        waves, valu = float(row[0]), float(row[1])
    except ValueError:
        return None
    insts = waves * valu
    achieved = insts / (ms * 1e-3) / 1e9
    return {"unit": "G VALU wave-instructions/s", "valu_per_wave": valu, "waves": int(waves),
            "achieved": round(achieved, 1), "peak": 614.4, "frac": round(achieved / 614.4, 4),
            "source": COUNTERS_PROFILE + " (SQ_WAVES, SQ_INSTS_VALU per dispatch)"}


def raster_kernel_name(kind: int, tile: int, waves: int) -> str:
    """The instantiated raster kernel: k_raster<tile, quadrants per wave, true>
    (kind 0, shading its own tiles) or k_raster_tp<tile, waves> (kind 1, which
    k_shade follows)."""
    return f"k_raster_tp<{tile}, {waves}>" if kind == 1 else f"k_raster<{tile}, {(tile // 8) ** 2 // waves}, true>"


def frame_rooflines(kind: int, kernel: str, P: int, entries: int, p_cov: int, t_vis: int, st: dict,
                    traffic: bool) -> dict:
    """The dominant kernel's roofline (the raster; for kind 0 it includes the
    shading), and for kind 1 the separate k_shade's."""
    tr = (lambda k: measured_traffic(k)) if traffic else (lambda k: (None, None))
    if kind == 0:
        return {"roofline": roofline(kernel, raster_shade_bytes(P, entries, p_cov, t_vis), st["ms_raster"],
                                     *tr(kernel), issue=traffic)}
    return {"roofline": roofline(kernel, raster_bytes(P, entries), st["ms_raster"], *tr(kernel), issue=traffic),
            "roofline_shade": roofline("k_shade", shade_bytes(P, p_cov, t_vis), st["ms_shade"], *tr("k_shade"),
                                       issue=traffic)}


def separate_shade_roofline(r, render, restore: tuple, P: int, p_cov: int, t_vis: int, traffic=(None, None),
                            issue: bool = False) -> dict:
    """k_shade on its own: the same frame (or batch) through the kind-1
    pipeline, whose raster writes the winner buffer and whose separate shading
    pass then reads it; the median HIP-event time of three eager launches.
    ``render`` must launch without the graph (a replay records no stages)."""
    r.set_raster(1, 32, 4)
    ms = []
    for _ in range(3):
        render()
        ms.append(r.stats()["ms_shade"])
    r.set_raster(*restore)
    return roofline("k_shade", shade_bytes(P, p_cov, t_vis), float(np.median(ms)), *traffic, issue=issue)


def roofline(kernel: str, nbytes: dict, ms: float, traffic=None, traffic_source=None, issue: bool = False) -> dict:
    """The HBM roofline of one kernel: algorithmic bytes over its HIP-event
    time against 8 TB/s.  ``traffic``: the bytes the kernel actually moves
    between L2 and memory per launch (committed counter summary, named in
    traffic_source), with the rate they imply.  ``issue``: add the VALU issue
    roofline from the committed counters, and name the limiter: the resource
    with the larger fraction of its peak."""
    gbs = nbytes["algorithmic"] / (ms * 1e-3) / 1e9 if ms > 0 else None
    out = {"kernel": kernel, "bound": "hbm", "achieved": round(gbs, 1) if gbs else None, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None, "traffic": traffic,
           "algorithmic_bytes": nbytes["algorithmic"], "record_bytes": nbytes["record"], "kernel_ms": round(ms, 4)}
    if traffic and ms > 0:
        tg = traffic / (ms * 1e-3) / 1e9
        out.update({"traffic_source": traffic_source, "traffic_gbs": round(tg, 1),
                    "traffic_frac": round(tg / HBM_PEAK_GBS, 4)})
    if issue:
        iss = issue_roofline(kernel, ms)
        if iss:
            out["issue"] = iss
            tf = out.get("traffic_frac") or 0.0
            out["limiter"] = ("VALU issue" if iss["frac"] >= tf else "memory traffic") + \
                f" ({max(iss['frac'], tf):.0%} of its peak)"
    return out


def cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo) and the usable core count."""
    name = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{name}, {len(os.sched_getaffinity(0))} usable cores"


def cpu_frames(scene, runs: int, warmup: int) -> dict:
    """The oracle (C restatement of the reference, one thread) rendering the
    whole frame: ``warmup`` untimed frames, then the median of ``runs`` timed
    ones (host wall clock around the full render)."""
    from oracle import oracle as O
    fr = O.frame_uniforms(scene)
    for _ in range(warmup):
        O.render(scene, frame=fr)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        r = O.render(scene, frame=fr)
        times.append(time.perf_counter() - t0)
        assert r.rc == 0
    covered = int(np.count_nonzero(r.winner != 0xFFFFFFFF))
    t = float(np.median(times))
    return {"frame_s": round(t, 4), "runs": runs, "warmup": warmup, "covered_px": covered,
            "shaded_mpix_per_s": round(covered / t / 1e6, 4), "tris_per_s": round(scene.triangle_count / t, 1)}


def cpu_baseline(scene, p_cov_full: int) -> dict:
    """The oracle (one thread) on one whole frame of the headline workload:
    the 1920x1080 1M-triangle soup, timed once (a frame takes seconds)."""
    c = cpu_frames(scene, runs=1, warmup=0)
    assert c["covered_px"] == p_cov_full, "the oracle's frame must cover what the GPU frame covers"
    return {
        "value": c["shaded_mpix_per_s"], "unit": "shaded Mpix/s", "cores": 1, "kind": "port",
        "sample": f"one whole 1920x1080 frame of the 1M-triangle soup (all {scene.triangle_count} input "
                  f"triangles), {c['frame_s']:.2f} s on {cpu_model()}",
        "tris_per_s": c["tris_per_s"], "frame_s": c["frame_s"],
    }


def winner_frame(r, render, W: int, H: int):
    """The winner ids of one more frame, rendered into a canvas that keeps
    them: the timed frames use canvases without a winner buffer
    (DH_CANVAS_NO_WINNER), as the reference's Canvas has none."""
    r.winners = True
    try:
        _, _, winner = r.read(render(), W, H)
    finally:
        r.winners = False
    return winner


SHELLS_WORKLOAD = ("20M micro-triangles on 8 nested jittered shells at 3840x2160 (5% flipped, 5% outside the "
                   "frustum), 8192x8192 texture, one dh_render per frame")


def bench_config(name: str, scene, steps: int, warmup: int, cpu_runs: int, cpu_warmup: int,
                 tiling=None) -> dict:
    """One more plain-path workload at its configured size: autotuned raster
    (or ``tiling`` = (kind, tile, waves)), ``steps`` timed frames (one
    dh_render each, graph-replayed), p50/p99 frame latency, and the oracle on
    whole frames of the same scene on this host."""
    from draw_amd.autotune import autotune
    from draw_amd.render import Renderer, frame_uniforms
    fr = frame_uniforms(scene)
    r = Renderer(0, profiling=True)
    r.winners = False
    meshes = r.mesh_handles(scene)
    if tiling is None:
        (kind, tile, waves), table = autotune(r, lambda: r.render_device(scene, fr, meshes))
    else:
        kind, tile, waves = tiling
        r.set_raster(kind, tile, waves)
    for _ in range(warmup):
        r.render_device(scene, fr, meshes)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        cv = r.render_device(scene, fr, meshes)
        lat.append(time.perf_counter() - a)
    total = time.perf_counter() - t0
    winner = winner_frame(r, lambda: r.render_device(scene, fr, meshes), fr.W, fr.H)
    p_cov, t_vis = visible_triangles(winner)
    st = r.stats()
    r.close()
    out = {"workload": name, "width": fr.W, "height": fr.H, "triangles": scene.triangle_count,
           "raster_kind": kind, "tile": tile, "waves_per_tile": waves, "steps": steps,
           "shaded_mpix_per_s": round(p_cov * steps / total / 1e6, 2),
           "tris_per_s": round(scene.triangle_count * steps / total, 1),
           "ms_per_frame": round(total / steps * 1e3, 4),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 4),
                          "p99": round(float(np.percentile(lat, 99)) * 1e3, 4)},
           "frame": {"covered_px": p_cov, "visible_tris": t_vis, "clipped_tris": int(st["tris_out"])}}
    if cpu_runs > 0:
        c = cpu_frames(scene, cpu_runs, cpu_warmup)
        assert c["covered_px"] == p_cov
        c["cores"] = 1
        c["host"] = cpu_model()
        out["cpu_oracle"] = c
        out["gpu_over_cpu"] = round(out["shaded_mpix_per_s"] / c["shaded_mpix_per_s"], 1)
    return out


def bench_batched(reps: int) -> dict:
    """256 views (8 seeded meshes of 1k-500k triangles x 32 cameras, resolutions
    uniform over 256x256 .. 1920x1080, 1024x1024 textures): the plain path
    (one dh_render per view, each complete before the next) against one
    batched launch replayed from a hipGraph.  Same views, same outputs."""
    from draw_amd import workloads
    from draw_amd.autotune import autotune
    from draw_amd.render import Renderer, frame_uniforms

    views = workloads.batched_scenes(seed=4)
    items = [(s, frame_uniforms(s)) for s in views]
    r = Renderer(0, profiling=True)
    r.winners = False  # as the reference's Canvas: frame and depth only
    prepared = r.batch_views(items)
    meshes = [r.mesh_handles(s) for s in views]
    log("batched: autotune ...")
    best, table = autotune(r, lambda: r.render_batch(items, prepared=prepared), frames=2)
    kind, tile, waves = best
    # plain path: one frame per launch
    for (s, fr), hs in zip(items[:8], meshes[:8]):
        r.render_device(s, fr, hs)
    lat = []
    t0 = time.perf_counter()
    for (s, fr), hs in zip(items, meshes):
        a = time.perf_counter()
        r.render_device(s, fr, hs)
        lat.append(time.perf_counter() - a)
    plain_s = time.perf_counter() - t0
    # batched: one launch for all views, replayed from a captured graph
    for _ in range(2):
        r.render_batch(items, graph=True, prepared=prepared)
    blat = []
    t0 = time.perf_counter()
    for _ in range(reps):
        a = time.perf_counter()
        cvs = r.render_batch(items, graph=True, prepared=prepared)
        blat.append(time.perf_counter() - a)
    batch_s = (time.perf_counter() - t0) / reps
    # one eager batch with per-stage HIP events (the graph replay records only the whole launch)
    r.render_batch(items, graph=False, prepared=prepared)
    st = r.stats()
    # covered pixels and winning triangles from one more batch into canvases
    # that keep the winner ids
    r.winners = True
    wcvs = r.render_batch(items, graph=False, prepared=r.batch_views(items, slot_base=len(items)))
    r.winners = False
    covered = t_vis = 0
    for cv, s in zip(wcvs, views):
        _, _, w = r.read(cv, s.width, s.height)
        pc, tv = visible_triangles(w)
        covered += pc
        t_vis += tv
    tris = sum(s.triangle_count for s in views)
    px = sum(s.width * s.height for s in views)
    log("batched: separate shading pass ...")
    shade_sep = separate_shade_roofline(r, lambda: r.render_batch(items, graph=False, prepared=prepared),
                                        (kind, tile, waves), px, covered, t_vis)
    rast_kernel = raster_kernel_name(kind, tile, waves)
    r.close()
    plain_mpix = covered / plain_s / 1e6
    batch_mpix = covered / batch_s / 1e6
    return {
        "workload": "256 views: 8 meshes (tori, UV spheres, soups; 1k-500k triangles log-uniform) x 32 random "
                    "cameras, resolutions uniform over 256x256/640x480/1280x720/1920x1080, 1024x1024 textures",
        "views": len(views), "triangles": tris, "pixels": px, "covered_px": covered,
        "raster_kind": kind, "tile": tile, "waves_per_tile": waves, "autotune_ms": table,
        "plain": {"shaded_mpix_per_s": round(plain_mpix, 1), "tris_per_s": round(tris / plain_s, 1),
                  "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 4),
                                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 4)},
                  "all_views_ms": round(plain_s * 1e3, 3)},
        "batched": {"shaded_mpix_per_s": round(batch_mpix, 1), "tris_per_s": round(tris / batch_s, 1),
                    "batch_latency_ms": {"p50": round(float(np.percentile(blat, 50)) * 1e3, 4),
                                         "p99": round(float(np.percentile(blat, 99)) * 1e3, 4)},
                    "all_views_ms": round(batch_s * 1e3, 3), "graph": True},
        "batched_over_plain": round(batch_mpix / plain_mpix, 3),
        "stage_ms_eager": {k: round(st[k], 4) for k in ("ms_geometry", "ms_binning", "ms_raster", "ms_shade",
                                                        "ms_frame")},
        **{("roofline_raster" if k == "roofline" else k): v
           for k, v in frame_rooflines(kind, rast_kernel, px, int(st["bin_entries"]), covered, t_vis, st,
                                       False).items()},
        "roofline_shade_separate": shade_sep,
    }


def bench_shard(scene, raster: tuple, steps: int = 5) -> dict:
    """One GPU's share of a sort-last frame at its real shape: the 1080p
    1M-triangle soup cut into 8 equal submission ranges (125k input
    triangles each), each shard rendered on its own context into its
    full-screen packed u64 visibility buffer (dh_render_shard), one u64-min
    merge of two such buffers (the per-buffer work of the cross-GPU
    reduction, done here on one device) and one shard's shading of the merged
    winners (dh_shade_shard).  Times: host wall clock around each synchronous
    call, median over ``steps``.  With 8 GPUs the frame waits for the slowest
    shard; the soup puts its near-plane straddlers first, so equal triangle
    counts are not equal work.  ``raster``: (kind, tile, waves) of the plain
    path's autotuned choice, used by every shard context."""
    from draw_amd import workloads
    from draw_amd.render import Renderer, frame_uniforms
    s8 = workloads.split_submission(scene, 8)
    W, H = s8.width, s8.height
    r = Renderer(0)
    r.set_raster(*raster)
    meshes = r.mesh_handles(s8)
    fr = frame_uniforms(s8)
    ctxs = [r.shard_context() for _ in range(8)]  # inherit the raster configuration
    cv = r.canvas(W, H)
    bufs = [r.device_alloc(W * H * 8) for _ in range(8)]
    med = lambda v: round(float(np.median(v)) * 1e3, 4)
    try:
        per_shard = []
        for k, cx in enumerate(ctxs):
            t = []
            for i in range(steps + 1):
                t0 = time.perf_counter()
                cx.render_shard(fr, [meshes[k]], k, cv, bufs[k])
                if i:
                    t.append(time.perf_counter() - t0)
            per_shard.append(med(t))
        t_m, t_s = [], []
        for i in range(steps + 1):
            t0 = time.perf_counter()
            r.vis_merge_min(bufs[0], bufs[1], W * H)
            t1 = time.perf_counter()
            ctxs[1].shade_shard(bufs[0], 1)
            if i:
                t_m.append(t1 - t0)
                t_s.append(time.perf_counter() - t1)
    finally:
        for p in bufs:
            r.device_free(p)
        for cx in ctxs:
            cx.close()
        r.close()
    return {"workload": "8 submission ranges of the 1920x1080 1M-triangle soup, 125k input triangles each",
            "raster_kind": raster[0], "tile": raster[1], "waves_per_tile": raster[2],
            "render_ms_per_shard": per_shard, "render_ms_max": max(per_shard),
            "merge_ms_per_buffer": med(t_m), "shade_ms_one_shard": med(t_s), "visibility_bytes": W * H * 8}


def bench_serving(seconds: float, rate_hz: float, n_distinct: int = 512, overload_seconds: float = 10.0) -> dict:
    """Config 5: a device-resident pool of 64 seeded meshes (1k-2M triangles,
    log-uniform, 1024^2 textures) and requests (mesh, random camera,
    resolution 854x480 p=.4 / 1280x720 .3 / 1920x1080 .25 / 3840x2160 .05).
    Three measurements over the same ``n_distinct`` requests:
      plain     one dh_render per request, back to back;
      capacity  a burst of 2 x n_distinct jobs through the server at once;
      stream    open-loop Poisson arrivals at ``rate_hz`` for ``seconds``
                (requests cycled), HBM budget 200 GB, p50/p99 latency from
                submit to launch completion;
      stream_3x_plain_rate  the same for ``overload_seconds`` at three times
                the request rate the plain path sustains, where arrivals pile
                up behind running launches and are batched.
    Shaded pixels per request come from one untimed render of each."""
    from draw_amd import workloads
    from draw_amd.render import Renderer, frame_uniforms
    from draw_amd.serve import Server, poisson_load

    log("serving: generating mesh pool ...")
    pool, scenes = workloads.serving_requests(seed=5, n_requests=n_distinct)
    r = Renderer(0, tile=32, waves=2, profiling=False)
    reqs = [(frame_uniforms(s), r.mesh_handles(s)) for s in scenes]
    pool_tris = int(sum(m.triangles.shape[0] for o in pool for m in o.meshes))
    log("serving: plain path ...")
    for fr, hs in reqs[:16]:
        r.render_device(None, fr, hs)
    t0 = time.perf_counter()
    for fr, hs in reqs:
        r.render_device(None, fr, hs)
    plain_s = time.perf_counter() - t0
    covered = []
    for fr, hs in reqs:
        cv = r.render_device(None, fr, hs)
        _, _, w = r.read(cv, fr.W, fr.H)
        covered.append(int(np.count_nonzero(w != 0xFFFFFFFF)))
    cov_total = sum(covered)
    tris_total = sum(s.triangle_count for s in scenes)
    srv = Server(r, budget_bytes=200 << 30, workers=2, max_views=64)
    log("serving: capacity burst ...")
    for fr, hs in reqs[:64]:
        srv.poll(srv.submit(fr, hs, autorelease=True), wait=True)
    t0 = time.perf_counter()
    tickets = [srv.submit(fr, hs, autorelease=True) for _ in range(2) for fr, hs in reqs]
    burst_lat = []
    for t in tickets:
        res = srv.poll(t, wait=True)
        burst_lat.append(res.latency_ms)
        srv.release(t)
    burst_s = time.perf_counter() - t0
    st0 = srv.stats()
    log(f"serving: Poisson {rate_hz:g}/s for {seconds:g} s ...")
    out = poisson_load(srv, reqs, rate_hz=rate_hz, seconds=seconds, seed=6)
    st1 = srv.stats()
    # the same stream at 3x the rate one dh_render per request can sustain:
    # jobs that arrive while launches run join the next launch
    plain_rps = len(reqs) / plain_s
    over_rate = round(3.0 * plain_rps, -2)
    over = None
    if overload_seconds > 0:
        log(f"serving: Poisson {over_rate:g}/s (3x the plain path) for {overload_seconds:g} s ...")
        st2 = srv.stats()
        over = poisson_load(srv, reqs, rate_hz=over_rate, seconds=overload_seconds, seed=7)
        st3 = srv.stats()
    srv.close()
    r.close()
    # served jobs cycle through the distinct requests in submission order
    served_cov = sum(covered[(res.ticket - st0["submitted"] - 1) % len(reqs)] for res, _ in out["results"])
    span = out["span_s"]
    overload = None
    if over is not None:
        o_cov = sum(covered[(res.ticket - st2["submitted"] - 1) % len(reqs)] for res, _ in over["results"])
        overload = {"rate_hz": over_rate, "seconds": overload_seconds, "requests": over["requests"],
                    "completed": over["completed"], "failed": over["failed"], "late_submits": over["late_submits"],
                    "latency_ms": over["latency_ms"], "mean_batch_views": over["mean_batch_views"],
                    "shaded_mpix_per_s": round(o_cov / over["span_s"] / 1e6, 1) if over["span_s"] > 0 else None,
                    "over_plain": round(o_cov / over["span_s"] / 1e6 / (cov_total / plain_s / 1e6), 3)
                    if over["span_s"] > 0 else None,
                    "launches": st3["launches"] - st2["launches"]}
    return {
        "workload": "64-mesh resident pool (1k-2M triangles log-uniform, 1024^2 RGBA8 textures), random cameras, "
                    "resolutions 854x480/1280x720/1920x1080/3840x2160 with p=.4/.3/.25/.05",
        "pool_triangles": pool_tris, "distinct_requests": len(reqs),
        "mean_tris_per_request": round(tris_total / len(reqs), 1),
        "mean_shaded_px_per_request": round(cov_total / len(reqs), 1),
        "plain": {"shaded_mpix_per_s": round(cov_total / plain_s / 1e6, 1),
                  "requests_per_s": round(len(reqs) / plain_s, 1), "tris_per_s": round(tris_total / plain_s, 1)},
        "capacity": {"shaded_mpix_per_s": round(2 * cov_total / burst_s / 1e6, 1),
                     "requests_per_s": round(2 * len(reqs) / burst_s, 1),
                     "tris_per_s": round(2 * tris_total / burst_s, 1),
                     "latency_ms": {"p50": round(float(np.percentile(burst_lat, 50)), 3),
                                    "p99": round(float(np.percentile(burst_lat, 99)), 3)}},
        "capacity_over_plain": round(plain_s * 2 / burst_s, 3),
        "stream": {"rate_hz": rate_hz, "seconds": seconds, "requests": out["requests"],
                   "completed": out["completed"], "failed": out["failed"], "late_submits": out["late_submits"],
                   "latency_ms": out["latency_ms"], "mean_batch_views": out["mean_batch_views"],
                   "shaded_mpix_per_s": round(served_cov / span / 1e6, 1) if span > 0 else None,
                   "launches": st1["launches"] - st0["launches"], "hbm_budget_bytes": 200 << 30,
                   "hbm_peak_bytes": st1["bytes_peak"]},
        "stream_3x_plain_rate": overload,
    }


def copy_calibration(mib: int) -> dict:
    """A plain device-to-device copy of ``mib`` MiB, three times: the known
    byte count that scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings."""
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(7)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    return {"copy_bytes": n, "copies": 3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--raster-kind", type=int, default=0, help="0: quadrant-owning waves, 1: triangle-parallel")
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--waves", type=int, default=4)
    ap.add_argument("--n-tri", type=int, default=1_000_000)
    ap.add_argument("--cpu", type=int, default=1, help="0 disables the CPU oracle timings")
    ap.add_argument("--configs", type=int, default=1,
                    help="0 skips the 640x480 torus and the 3840x2160 20M-triangle legs; 2 runs only the "
                         "20M-triangle leg (for profiling it alone)")
    ap.add_argument("--batched-reps", type=int, default=5, help="timed batched launches; 0 skips the batched run")
    ap.add_argument("--shard-steps", type=int, default=5,
                    help="timed calls of the 1-of-8 sort-last shard leg; 0 skips it")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the Poisson serving stream (60 s, as the serving config names); 0 skips serving")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrival rate of the serving stream")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="use --tile/--waves instead of the autotuner's choice")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (calibrates rocprofv3 byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        print(json.dumps(copy_calibration(args.copy_calibration)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for {args.gpus}); the sort-last shard leg times one GPU's "
              f"share of an 8-way split, but no multi-GPU run is made", file=sys.stderr)
        sys.exit(2)

    from draw_amd import workloads
    from draw_amd.render import Renderer, frame_uniforms

    if args.configs == 2:
        log("shells 3840x2160, 20M triangles only ...")
        print(json.dumps({"shells_20m_4k": bench_config(
            SHELLS_WORKLOAD, workloads.shells_scene(3), steps=args.steps, warmup=args.warmup,
            cpu_runs=1 if args.cpu else 0, cpu_warmup=0,
            tiling=None if args.autotune else (args.raster_kind, args.tile, args.waves))}), flush=True)
        return

    log("generating workload ...")
    scene = workloads.soup_scene(2, n_tri=args.n_tri)
    fr = frame_uniforms(scene)
    W, H = fr.W, fr.H
    r = Renderer(0, profiling=True)
    r.winners = False  # as the reference's Canvas: frame and depth only
    r.set_raster(args.raster_kind, args.tile, args.waves)
    meshes = r.mesh_handles(scene)
    tune_table = None
    if args.autotune:
        from draw_amd.autotune import autotune
        log("autotuning raster kernel / tile size / waves per tile ...")
        (args.raster_kind, args.tile, args.waves), tune_table = autotune(r, lambda: r.render_device(scene, fr, meshes))
        log(f"autotune: {tune_table} -> kind {args.raster_kind} {args.tile}x{args.waves}")
    log("uploaded; warmup ...")
    for _ in range(args.warmup):
        r.render_device(scene, fr, meshes)
    lat, graph_ms = [], []
    log(f"timing {args.steps} frames ...")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        cv = r.render_device(scene, fr, meshes)
        lat.append(time.perf_counter() - a)
        graph_ms.append(r.stats()["ms_frame"])
    total = time.perf_counter() - t0
    winner = winner_frame(r, lambda: r.render_device(scene, fr, meshes), W, H)
    p_cov, t_vis = visible_triangles(winner)
    # the drop-in's whole Scene::render: the frame plus dh_canvas_read of the
    # BGRA8 frame and f32 depth into caller-owned (pageable) host buffers
    host_bgra = np.empty((H, W, 4), np.uint8)
    host_depth = np.empty((H, W), np.float32)
    rb = []
    for _ in range(max(5, args.steps // 2)):
        a = time.perf_counter()
        cv = r.render_device(scene, fr, meshes)
        r.read_into(cv, host_bgra, host_depth)
        rb.append(time.perf_counter() - a)
    rb_ms = float(np.median(rb)) * 1e3
    # per-stage HIP event times need eager launches (a replayed graph records
    # only the whole frame): the same frame again, untimed, without the graph
    r.set_graph(False)
    st_acc = []
    for _ in range(max(3, min(args.steps, 10))):
        r.render_device(scene, fr, meshes)
        st_acc.append(r.stats())
    st = {k: float(np.mean([s[k] for s in st_acc])) for k in st_acc[0]}
    shade_sep = None
    if args.raster_kind == 0:
        shade_sep = separate_shade_roofline(r, lambda: r.render_device(scene, fr, meshes),
                                            (args.raster_kind, args.tile, args.waves), W * H, p_cov, t_vis,
                                            measured_traffic("k_shade"), issue=True)
    r.set_graph(True)
    frame_ms = total / args.steps * 1e3
    value = p_cov * args.steps / total / 1e6
    rast_kernel = raster_kernel_name(args.raster_kind, args.tile, args.waves)
    roofs = frame_rooflines(args.raster_kind, rast_kernel, W * H, int(st["bin_entries"]), p_cov, t_vis, st, True)
    out = {
        "metric": "shaded Mpix/s", "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(frame_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain GPU path, 1920x1080, seeded 1M-triangle soup (20% near-plane straddlers, "
                               "log-normal area median 40 px, random winding: half the triangles face away and "
                               "are culled), 2048x2048 RGBA8 texture, one frame per dh_render into a canvas "
                               "resident in HBM (BGRA8 and f32 depth, no winner buffer, as the reference's "
                               "Canvas; launch sequence replayed from a hipGraph)",
                   "width": W, "height": H, "triangles": scene.triangle_count, "back_facing_frac": 0.5,
                   "raster_kind": args.raster_kind,
                   "tile": args.tile,
                   "waves_per_tile": args.waves, "autotune_ms": tune_table},
        "tris_per_s": round(scene.triangle_count * args.steps / total, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 4),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 4)},
        "device_ms_per_frame": round(float(np.mean(graph_ms)), 4),
        "readback_inclusive": {"what": "dh_render + dh_canvas_read of BGRA8 and f32 depth (16.6 MB) into "
                                       "pageable host memory, median per frame",
                               "ms_per_frame": round(rb_ms, 4),
                               "shaded_mpix_per_s": round(p_cov / rb_ms / 1e3, 2)},
        "stage_ms_eager": {k: round(st[k], 4) for k in ("ms_geometry", "ms_binning", "ms_raster", "ms_shade",
                                                        "ms_frame")},
        "frame": {"covered_px": p_cov, "visible_tris": t_vis, "clipped_tris": int(st["tris_out"]),
                  "bin_entries": int(st["bin_entries"]), "fragments": int(st["frags"]),
                  "frags_per_px": round(st["frags"] / (W * H), 2)},
        **roofs,
    }
    if shade_sep is not None:
        out["roofline_shade"] = shade_sep
    r.close()
    if args.shard_steps > 0:
        log("sort-last shard (1 of 8) ...")
        out["shard_1_of_8"] = bench_shard(scene, (args.raster_kind, args.tile, args.waves), args.shard_steps)
    if args.batched_reps > 0:
        out["batched"] = bench_batched(args.batched_reps)
    if args.serve_seconds > 0:
        out["serving"] = bench_serving(args.serve_seconds, args.serve_rate)
    if args.configs > 0:
        log("torus 640x480 ...")
        out["torus_640x480"] = bench_config(
            "2304-triangle torus (R 1.0, r 0.4, 48x24) at 640x480, 256x256 checker, one dh_render per frame",
            workloads.torus_scene(1), steps=200, warmup=20, cpu_runs=5 if args.cpu else 0, cpu_warmup=1)
        log("shells 3840x2160, 20M triangles ...")
        out["shells_20m_4k"] = bench_config(SHELLS_WORKLOAD, workloads.shells_scene(3), steps=20, warmup=3,
            cpu_runs=1 if args.cpu else 0, cpu_warmup=0)
    if args.cpu > 0:
        log("CPU baseline: one whole frame of the soup ...")
        out["cpu_baseline"] = cpu_baseline(scene, p_cov)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
