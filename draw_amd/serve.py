"""Render serving on libdrawhip: the continuous-batching server of the C ABI
(dh_srv_*) and a Poisson load generator.

``Server`` wraps one dh_server over a ``Renderer``'s context: meshes and
textures are uploaded once through the renderer (resident in HBM) and jobs
reference them.  Each job is one Scene::render (scene/mod.rs:901) of its
meshes with its own camera and resolution; the server batches whatever is
queued when a worker frees up, within an HBM budget.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import DhFrameParams, DhJob, DhJobInfo, DhServerConfig, DhServerStats, DrawHipError
from .render import Renderer


@dataclass
class JobResult:
    ticket: int
    status: int
    canvas: int
    batch_views: int
    t_submit_ms: float
    t_admit_ms: float
    t_done_ms: float

    @property
    def latency_ms(self) -> float:
        return self.t_done_ms - self.t_submit_ms


class Server:
    def __init__(self, renderer: Renderer, budget_bytes: int = 0, workers: int = 2, max_views: int = 64,
                 max_tris: int = 16 << 20):
        self.r = renderer
        self.L = renderer.L
        cfg = DhServerConfig(int(budget_bytes), int(workers), int(max_views), int(max_tris))
        srv = C.c_void_p()
        rc = self.L.dh_srv_create(renderer.ctx, C.byref(cfg), C.byref(srv))
        if rc != 0:
            raise DrawHipError(rc, "dh_srv_create")
        self.srv = srv
        self._keep = {}

    def submit(self, frame: DhFrameParams, meshes: List[int], autorelease: bool = False) -> int:
        arr = (C.c_uint64 * max(1, len(meshes)))(*meshes)
        job = DhJob(frame, C.cast(arr, C.POINTER(C.c_uint64)), len(meshes),
                    _lib.DH_JOB_AUTORELEASE if autorelease else 0)
        t = C.c_uint64()
        rc = self.L.dh_srv_submit(self.srv, C.byref(job), C.byref(t))
        if rc != 0:
            raise DrawHipError(rc, "dh_srv_submit")
        return t.value

    def poll(self, ticket: int, wait: bool = True) -> JobResult:
        info = DhJobInfo()
        rc = self.L.dh_srv_poll(self.srv, int(ticket), int(bool(wait)), C.byref(info))
        if rc != 0:
            raise DrawHipError(rc, "dh_srv_poll")
        return JobResult(ticket, info.status, info.canvas, info.batch_views, info.t_submit_ms, info.t_admit_ms,
                         info.t_done_ms)

    def release(self, ticket: int):
        rc = self.L.dh_srv_release(self.srv, int(ticket))
        if rc != 0:
            raise DrawHipError(rc, "dh_srv_release")

    def stats(self) -> dict:
        st = DhServerStats()
        rc = self.L.dh_srv_stats(self.srv, C.byref(st))
        if rc != 0:
            raise DrawHipError(rc, "dh_srv_stats")
        return {n: getattr(st, n) for n, _ in DhServerStats._fields_}

    def read(self, result: JobResult, W: int, H: int):
        return self.r.read(result.canvas, W, H)

    def close(self):
        if getattr(self, "srv", None):
            self.L.dh_srv_destroy(self.srv)
            self.srv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_load(server: Server, requests, rate_hz: float, seconds: float, seed: int = 0,
                 collector_threads: int = 2) -> dict:
    """Open-loop load: arrival times are a Poisson process of ``rate_hz`` over
    ``seconds``; each arrival submits the next request (frame, meshes) from
    ``requests`` (cycled) with autorelease, on schedule whether or not earlier
    jobs finished.  Collector threads wait for completions.  Returns per-job
    latencies (server clock: submit -> launch complete), batch sizes and
    throughput."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, size=int(rate_hz * seconds * 1.5) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < seconds]
    tickets: List[tuple] = []
    lock = threading.Lock()
    done = []
    submitted_all = threading.Event()

    def collect(k0):
        i = k0
        while True:
            with lock:
                n = len(tickets)
            if i >= n:
                if submitted_all.is_set():
                    with lock:
                        if i >= len(tickets):
                            return
                    continue
                time.sleep(0.0005)
                continue
            t, px = tickets[i]
            res = server.poll(t, wait=True)
            server.release(t)
            with lock:
                done.append((res, px))
            i += collector_threads

    ths = [threading.Thread(target=collect, args=(k,), daemon=True) for k in range(collector_threads)]
    for th in ths:
        th.start()
    t0 = time.perf_counter()
    late = 0
    for k, ta in enumerate(arrivals):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        elif now - ta > 0.005:
            late += 1
        fr, meshes = requests[k % len(requests)]
        t = server.submit(fr, meshes, autorelease=True)
        with lock:
            tickets.append((t, fr.W * fr.H))
    submitted_all.set()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    ok = [(r, px) for r, px in done if r.status == 0]
    lat = np.array([r.latency_ms for r, _ in ok]) if ok else np.zeros(1)
    t_first = min(r.t_submit_ms for r, _ in ok) if ok else 0.0
    t_last = max(r.t_done_ms for r, _ in ok) if ok else 1.0
    return {
        "requests": len(arrivals), "completed": len(ok), "failed": len(done) - len(ok),
        "late_submits": late, "wall_s": round(wall, 3),
        "span_s": round((t_last - t_first) / 1e3, 3),
        "pixels": int(sum(px for _, px in ok)),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "mean": round(float(lat.mean()), 3), "max": round(float(lat.max()), 3)},
        "mean_batch_views": round(float(np.mean([r.batch_views for r, _ in ok])), 2) if ok else 0.0,
        "results": ok,
    }
