"""Served form of the AVBD step: the request server of the C ABI
(avbd_server_*, include/avbd_capi.h) and an open-loop Poisson load generator.

The scheduler itself is native (skyephyx_amd/csrc/avbd_server.cpp): scenes
resident on the device under an HBM byte budget, requests admitted in arrival
order into free slots, one batched device step per tick over every in-flight
scene.  This module binds it and drives it with arrivals.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .abi import AvbdRequestInfo, AvbdServerStats, AvbdWorld, alloc_world, world_to_c
from .capi import Context, _check, lib
from .world import World

STATE_QUEUED, STATE_RUNNING, STATE_DONE = 0, 1, 2


def _info(r: AvbdRequestInfo) -> dict:
    return {name: getattr(r, name) for name, _ in AvbdRequestInfo._fields_ if name != "pad_"}


class Server:
    """Request server on one GPU: budget in bytes of device memory for the
    resident scenes, at most max_slots requests in flight."""

    def __init__(self, budget_bytes: float = 200e9, max_slots: int = 1 << 20, device: int = 0,
                 lanes: int | None = None, fem_lanes: int | None = None):
        self.L = lib()
        self.h = C.c_void_p()
        _check(self.L.avbd_server_create(int(device), int(budget_bytes), int(max_slots), C.byref(self.h)))
        self.context = Context(lanes=lanes, fem_lanes=fem_lanes, _handle=self.L.avbd_server_context(self.h))
        self.n_scenes = 0

    def close(self):
        if self.h:
            self.context.close()
            self.L.avbd_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_scenes(self, world: World) -> range:
        """Register every scene of `world`; returns their ids."""
        keep, cw = world_to_c(world)
        first = C.c_int32()
        _check(self.L.avbd_server_add_scenes(self.h, C.byref(cw), C.byref(first)))
        self.n_scenes = first.value + len(world.scenes)
        return range(first.value, first.value + len(world.scenes))

    def preload(self):
        _check(self.L.avbd_server_preload(self.h))

    def now(self) -> float:
        return float(self.L.avbd_server_now(self.h))

    def submit(self, scene: int, steps: int, t_arrival: float = -1.0) -> int:
        t = C.c_int64()
        _check(self.L.avbd_server_submit(self.h, int(scene), int(steps), float(t_arrival), C.byref(t)))
        return int(t.value)

    def tick(self) -> int:
        n = C.c_int32()
        _check(self.L.avbd_server_tick(self.h, C.byref(n)))
        return int(n.value)

    def request(self, ticket: int) -> dict:
        r = AvbdRequestInfo()
        _check(self.L.avbd_server_request(self.h, int(ticket), C.byref(r)))
        return _info(r)

    def poll(self, cap: int = 65536) -> list[dict]:
        """Finished requests since the last poll, in completion order."""
        out = []
        buf = (AvbdRequestInfo * cap)()
        while True:
            n = C.c_int32()
            _check(self.L.avbd_server_poll(self.h, buf, cap, C.byref(n)))
            out.extend(_info(buf[i]) for i in range(n.value))
            if n.value < cap:
                return out

    def stats(self) -> dict:
        s = AvbdServerStats()
        _check(self.L.avbd_server_get_stats(self.h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in AvbdServerStats._fields_ if name != "pad_"}

    def download_scene(self, scene: int) -> World:
        cnt = AvbdWorld()
        _check(self.L.avbd_server_download_scene(self.h, int(scene), C.byref(cnt)))
        w, cw, keep = alloc_world(cnt)
        _check(self.L.avbd_server_download_scene(self.h, int(scene), C.byref(cw)))
        return w

    def run_until_idle(self, max_ticks: int = 1 << 30) -> int:
        """Tick until nothing is queued or running; returns the ticks taken."""
        k = 0
        while k < max_ticks:
            s = self.stats()
            if s["queued"] == 0 and s["running"] == 0:
                break
            self.tick()
            k += 1
        return k


def _pct(x, q):
    return round(float(np.percentile(x, q)), 3) if len(x) else None


def _drain(server: Server, limit_s: float):
    """Tick until nothing is queued or running, for at most limit_s seconds."""
    t0 = server.now()
    while server.now() - t0 < limit_s:
        st = server.stats()
        if st["queued"] == 0 and st["running"] == 0:
            return
        server.tick()


def closed_loop(server: Server, scene_ids, duration: float, steps=(1, 8), seed: int = 0) -> dict:
    """Saturation of the served form: every scene always has a request in
    flight (a finished request is replaced at once by a new one for the same
    scene, U{steps} steps), for `duration` seconds.  Returns the served
    scene-steps/s and the tick time."""
    rng = np.random.default_rng(seed)
    ids = [int(s) for s in scene_ids]
    _drain(server, 600.0)
    server.poll()
    for s in ids:
        server.submit(s, int(rng.integers(steps[0], steps[1] + 1)))
    s0 = server.stats()
    t0 = server.now()
    tick_ms, width = [], []
    while server.now() - t0 < duration:
        a = server.now()
        width.append(server.tick())
        tick_ms.append((server.now() - a) * 1e3)
        for d in server.poll():
            server.submit(d["scene"], int(rng.integers(steps[0], steps[1] + 1)))
    elapsed = server.now() - t0
    s1 = server.stats()
    _drain(server, 600.0)
    server.poll()
    return {
        "scene_steps_per_s": round((s1["scene_steps"] - s0["scene_steps"]) / elapsed, 1),
        "tick_ms_p50": _pct(tick_ms, 50),
        "scenes_per_tick_mean": round(float(np.mean(width)), 1) if width else 0.0,
        "elapsed_s": round(elapsed, 3),
    }


def poisson_run(server: Server, scene_ids, rate: float, duration: float, steps=(1, 8), seed: int = 0,
                log=None, drain_limit: float | None = None) -> dict:
    """Open-loop load: requests arrive as a Poisson process of `rate` per
    second for `duration` seconds, each for a uniformly drawn scene and
    U{steps[0]..steps[1]} steps.  A request's arrival time is its scheduled
    time, so time spent waiting for a tick to end counts as queueing.  After
    the arrivals the server drains for at most drain_limit seconds (default
    2 x duration); requests still queued or running then are reported as the
    backlog (an overloaded server never drains).  Returns latency percentiles
    (ms) of the finished requests and the served scene-steps/s over the
    arrival window."""
    rng = np.random.default_rng(seed)
    n_max = int(rate * duration * 1.5) + 64
    arr = np.cumsum(rng.exponential(1.0 / rate, n_max))
    arr = arr[arr < duration]
    sc = rng.choice(np.asarray(scene_ids), len(arr))
    ks = rng.integers(steps[0], steps[1] + 1, len(arr))
    drain_limit = 2.0 * duration if drain_limit is None else drain_limit
    server.poll()
    s0 = server.stats()
    t0 = server.now()
    i, tick_ms, width = 0, [], []
    next_log = 5.0
    s_window = None
    while True:
        now = server.now() - t0
        while i < len(arr) and arr[i] <= now:
            server.submit(int(sc[i]), int(ks[i]), t0 + float(arr[i]))
            i += 1
        if s_window is None and now >= duration:
            s_window = server.stats()
        st = server.stats()
        if st["queued"] == 0 and st["running"] == 0:
            if i >= len(arr):
                break
            time.sleep(max(0.0, float(arr[i]) - (server.now() - t0)))
            continue
        if now > duration + drain_limit:
            break
        a = server.now()
        n = server.tick()
        tick_ms.append((server.now() - a) * 1e3)
        width.append(n)
        if log is not None and now > next_log:
            log(f"[serve] t={now:.1f}s submitted {i}/{len(arr)} queued {st['queued']} running {st['running']}")
            next_log += 5.0
    if s_window is None:
        s_window = server.stats()
    done = server.poll()
    end = server.stats()
    backlog = end["queued"] + end["running"]
    _drain(server, 600.0)  # leave the server idle for the next run
    server.poll()
    lat = np.array([(d["t_done"] - d["t_arrival"]) * 1e3 for d in done])
    wait = np.array([(d["t_start"] - d["t_arrival"]) * 1e3 for d in done])
    steps_window = s_window["scene_steps"] - s0["scene_steps"]
    return {
        "rate_req_per_s": round(rate, 2),
        "offered_scene_steps_per_s": round(rate * (steps[0] + steps[1]) / 2.0, 1),
        "requests": int(len(arr)),
        "finished": int(len(done)),
        "backlog": int(backlog),
        "served_scene_steps_per_s": round(steps_window / duration, 1),
        "latency_ms_p50": _pct(lat, 50),
        "latency_ms_p99": _pct(lat, 99),
        "queue_ms_p50": _pct(wait, 50),
        "queue_ms_p99": _pct(wait, 99),
        "tick_ms_p50": _pct(tick_ms, 50),
        "tick_ms_p99": _pct(tick_ms, 99),
        "scenes_per_tick_mean": round(float(np.mean(width)), 1) if width else 0.0,
        "ticks": int(len(tick_ms)),
    }
