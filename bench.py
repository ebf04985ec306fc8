#!/usr/bin/env python3
"""Benchmark of the MI355X AVBD step.

Default workload: the batched mix (4096 independent scenes of uneven size:
rigid box piles with 8-2048 boxes, hanging Neo-Hookean blocks with 1k-50k
tets, per-scene dt in {1/60, 1/120} and 8-16 iterations), resident in HBM and
advanced with one step call over all scenes.  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mix|fem|boxes20k|pyramid]

After the upload the world settles --settle steps (default 150) before the
W warm-up steps, so the timed steps carry the steady contact load (the box
piles roughly triple their contacts over the first ~100 steps) whatever W
is.  The K timed steps run as the product does (step graphs, no per-launch
events).

value = whole-job scene-steps/s.  The line also carries the roofline of the
dominant kernel (the primal block-solve sweep) and of the other sweep
kernels, from an instrumented pass of --roofline-steps steps after the
timed ones (HIP events around every launch on the stream it runs on), the
served form over the same scenes (request server with every scene resident
under a 200 GB HBM budget, Poisson arrivals at --serve-load fractions of its
own closed-loop saturation: request latency p50/p99 and tick p99), a CPU
baseline (the FP32 oracle, a restatement of the reference's single-threaded
Solver::Step, timed on one host core over a bounded, size-stratified sample
of the same scenes) and the plain GPU step against that CPU step on the
landed 20k-box world.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "batched scene-steps/s + p99 step latency; body/vertex updates/s vs plain/CPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name):
    from skyephyx_amd import scenes, scenes_fast

    if name == "mix":
        return scenes_fast.batched_mix(4096), "batched mix, 4096 scenes (boxes 8-2048 log-uniform, FEM 1k-50k tets)"
    if name == "fem":
        return scenes_fast.fem_block(32, cell=0.25), "FEM block 32^3 vertices, 148955 tets, Neo-Hookean E=1e5 nu=0.3"
    if name == "boxes20k":
        return (scenes.box_lattice(shape=(40, 25, 20), spacing=3.5, n_chain=1000, n_springs=5000),
                "20k boxes in a 40x25x20 lattice + 1k-link ball-joint chain + 5k springs")
    if name == "pyramid":
        return scenes.pyramid_2d(), "2-D pyramid of 55 unit boxes on a static ground"
    raise SystemExit(f"unknown workload {name}")


KERNEL_CLASSES = ("primal_forces", "primal_fem", "dual")
KERNEL_NAMES = {
    "primal_fem": "k_primal_fem<G> (FEM-only bucket of a primal colour sweep)",
    "primal_forces": "k_primal<G, false> (bucket of bodies with forces)",
    "dual": "k_dual (dual and penalty update of every contact, joint and spring row)",
}


def measured_traffic(workload, lanes, cls):
    """HBM bytes per launch of the kernel class from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known
    size), or None when no summary exists for this workload (autotuned
    lanes only)."""
    tag = "k_primal_fem" if cls == "primal_fem" else "k_primal_forces"
    path = os.path.join(ROOT, "profiles", f"{tag}_batched_{workload}_hbm.md")
    if lanes != 0 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def stratified_sample(world, k, seed=3):
    """k scene ids, one from each size quantile bin."""
    sizes = np.bincount(world.bodies["scene"], minlength=len(world.scenes)).astype(np.float64)
    if len(world.tets):
        sizes += np.bincount(world.bodies["scene"][world.tets["v"][:, 0]], minlength=len(world.scenes))
    order = np.argsort(sizes, kind="stable")
    bins = np.array_split(order, min(k, len(order)))
    rng = np.random.default_rng(seed)
    return [int(rng.choice(b)) for b in bins if len(b)], [len(b) for b in bins if len(b)]


def cpu_baseline(world, budget_s=20.0):
    """The FP32 oracle (a restatement of the reference's single-threaded
    Solver::Step) on one host core, over a size-stratified sample of the
    scenes (all of them if there is one): per scene up to 10 untimed warm-up
    steps, then steps timed one by one, at least 3 and as many as the
    scene's share of the budget allows (at most 300); the median step time
    per scene, weighted by the scenes its size bin stands for."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    S = len(world.scenes)
    ids, weights = stratified_sample(world, 16) if S > 1 else ([0], [1])
    share = budget_s / len(ids)
    total_w, t_w, spent, n_timed = 0, 0.0, 0.0, 0
    per = []
    for sid, wgt in zip(ids, weights):
        sub = world.extract_scene(sid) if S > 1 else world
        o = O.Oracle(sub, fp64=False, order=O.NATURAL)
        t0 = time.perf_counter()
        o.step(1)
        first = time.perf_counter() - t0
        warm = min(9, int(0.2 * share / max(first, 1e-6)))
        if warm > 0:
            o.step(warm)
        times = []
        while len(times) < 3 or (sum(times) < 0.8 * share and len(times) < 300):
            t0 = time.perf_counter()
            o.step(1)
            times.append(time.perf_counter() - t0)
        med = float(np.median(times))
        spent += first * (1 + warm) + sum(times)
        n_timed += len(times)
        total_w += wgt
        t_w += wgt * med
        per.append((sid, len(sub.bodies), len(sub.tets), med))
    mean_t = t_w / total_w  # estimated mean CPU seconds per scene-step
    desc = (f"{len(ids)} scene(s) (one per size bin of {S}) from the GPU's post-benchmark state, up to 10 warm-up "
            f"steps each, then {n_timed} steps timed one by one (median per scene); {spent:.1f} s of CPU work; "
            f"FP32 oracle, natural order, 1 thread")
    return 1.0 / mean_t, desc, per, (ids, weights)


def cpu_all_cores(world, ids, weights, seconds=8.0):
    """Independent scenes on all host cores (one scene per thread, the
    sampled scenes spread over the threads; the oracle's step releases the
    GIL): the weighted median step time under that contention gives the
    whole-host scene-steps/s."""
    import threading

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    threads = max(1, min(16, os.cpu_count() or 1, len(ids)))
    oracles = [O.Oracle(world.extract_scene(s), fp64=False, order=O.NATURAL) for s in ids]
    times = [[] for _ in ids]
    stop = time.perf_counter() + seconds

    def work(k):
        mine = list(range(k, len(ids), threads))
        while time.perf_counter() < stop:
            for i in mine:
                t0 = time.perf_counter()
                oracles[i].step(1)
                times[i].append(time.perf_counter() - t0)
                if time.perf_counter() >= stop:
                    break

    pool = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    t_w = sum(w * float(np.median(t)) for w, t in zip(weights, times) if t)
    w_sum = sum(w for w, t in zip(weights, times) if t)
    # each thread advances its scenes in turn: threads scene-steps per mean step time
    return threads / (t_w / w_sum), threads, sum(len(t) for t in times)


def plain_gpu_rate(world, steps=10, configs=((8, 16), (4, 4))):
    """Plain per-scene GPU step: each sampled scene alone in its own context,
    at the faster of two lane settings (8/16 lanes suit large scenes, 4/4
    small ones)."""
    from skyephyx_amd.capi import Context

    ids, weights = stratified_sample(world, 16, seed=5)
    t_w, total_w = 0.0, 0
    for sid, wgt in zip(ids, weights):
        sub = world.extract_scene(sid)
        best = math.inf
        for lanes, fem_lanes in configs:
            with Context(lanes=lanes, fem_lanes=fem_lanes) as ctx:
                ctx.upload(sub)
                ctx.step(2)
                ctx.synchronize()
                t0 = time.perf_counter()
                ctx.step(steps)
                ctx.synchronize()
                best = min(best, (time.perf_counter() - t0) / steps)
        t_w += wgt * best
        total_w += wgt
    return 1.0 / (t_w / total_w)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def boxes20k_plain_vs_cpu(no_cpu, settle=360, steps=30, cpu_steps=3):
    """The plain per-scene GPU step against the CPU step on the 20k-box
    world (40 x 25 x 20 lattice, 1k-link chain, 5k springs) once its boxes
    have landed: the GPU world settles `settle` steps, then `steps` steps
    are timed (step graphs); the landed state is downloaded and the FP32
    oracle (one thread, natural order) times `cpu_steps` steps of it after
    one untimed step."""
    from skyephyx_amd import scenes
    from skyephyx_amd.capi import Context

    world = scenes.box_lattice(shape=(40, 25, 20), spacing=3.5, n_chain=1000, n_springs=5000)
    with Context() as ctx:
        ctx.upload(world)
        for i in range(0, settle, 30):
            ctx.step(min(30, settle - i))
        ctx.step(5)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.synchronize()
        gpu = steps / (time.perf_counter() - t0)
        state = ctx.download()
        em, _ = ctx.status(1)
    out = {"gpu_steps_per_s": round(gpu, 2), "state": f"after {settle + 5} GPU steps: {len(state.manifolds)} "
           f"manifolds, {len(state.bodies)} bodies, emergency stop {int(em[0])}", "gpu_steps_timed": steps}
    log(f"[bench]   20k boxes GPU {gpu:.1f} steps/s ({out['state']})")
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O

        o = O.Oracle(state, fp64=False, order=O.NATURAL)
        o.step(1)
        times = []
        for _ in range(cpu_steps):
            t0 = time.perf_counter()
            o.step(1)
            times.append(time.perf_counter() - t0)
        cpu = 1.0 / float(np.median(times))
        out.update(cpu_steps_per_s=round(cpu, 4), cpu_cores=1, cpu_kind="port",
                   cpu_sample=f"FP32 oracle, 1 thread, {cpu_steps} steps timed after 1 untimed, median",
                   gpu_over_cpu=round(gpu / cpu, 1))
        log(f"[bench]   20k boxes CPU {cpu:.3f} steps/s -> {gpu / cpu:.0f}x")
    return out


def served(world, fractions, duration, budget=200e9):
    """Served form over the same scenes: every scene registered and resident
    under an HBM budget, requests of U{1..8} steps of a random scene.  The
    served saturation is measured first (closed loop: every scene always has
    a request in flight), then Poisson arrivals at fractions of it.  The
    server keeps its default pinned lane widths (a scene's result does not
    depend on the other scenes in its ticks)."""
    from skyephyx_amd.serve import Server, closed_loop, poisson_run

    mean_k = 4.5
    out = []
    with Server(budget_bytes=budget) as srv:
        t0 = time.perf_counter()
        ids = srv.add_scenes(world)
        srv.preload()
        st = srv.stats()
        log(f"[serve] {len(ids)} scenes registered and resident ({st['resident_bytes'] / 1e9:.2f} GB of "
            f"{budget / 1e9:.0f} GB budget) in {time.perf_counter() - t0:.1f}s")
        closed_loop(srv, ids, 1.0, seed=98)  # untimed warm-up
        sat = closed_loop(srv, ids, 4.0, seed=99)
        log(f"[serve] saturation (closed loop): {json.dumps(sat)}")
        for f in fractions:
            rate = f * sat["scene_steps_per_s"] / mean_k
            r = poisson_run(srv, ids, rate, duration, steps=(1, 8), seed=int(f * 1000), log=log)
            r["load_fraction"] = f
            log(f"[serve] load {f:.2f}: {json.dumps(r)}")
            out.append(r)
        timings = srv.context.timings()
        st = srv.stats()
    return out, st, sat, timings


COPY_CALIBRATION_BYTES = 512 * 1024 * 1024


def copy_calibration():
    """One device copy of COPY_CALIBRATION_BYTES (read once, written once)."""
    import torch

    x = torch.ones(COPY_CALIBRATION_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {COPY_CALIBRATION_BYTES} bytes", file=sys.stderr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle", type=int, default=150,
                    help="steps run after the upload and before the warm-up, so the timed steps see the settled "
                         "contact load whatever --warmup is (the mix's piles roughly triple their contacts over "
                         "the first ~100 steps)")
    ap.add_argument("--roofline-steps", type=int, default=20,
                    help="steps of the instrumented pass after the timed steps (eager launches with an event "
                         "pair around every sweep launch; the source of the roofline figures)")
    ap.add_argument("--workload", default="mix")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per body with forces (0 = autotuned)")
    ap.add_argument("--fem-lanes", type=int, default=0, help="lanes per FEM-only body (0 = autotuned)")
    ap.add_argument("--no-roofline", action="store_true", help="skip the instrumented pass (no roofline figures)")
    ap.add_argument("--no-boxes20k", action="store_true",
                    help="skip the plain GPU step vs the CPU step on the landed 20k-box world")
    ap.add_argument("--no-graphs", action="store_true", help="launch every kernel of a step one by one")
    ap.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                    help="avbd_set_option on the benchmark context (repeatable)")
    ap.add_argument("--fused", action="store_true", help="small scenes in the per-scene kernel (option fused_scenes)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-scene comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the served (Poisson request) measurement")
    ap.add_argument("--serve-load", type=float, nargs="+", default=[0.1, 0.5, 0.8, 1.0],
                    help="served load points as fractions of the served (closed-loop) saturation")
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="arrival window per served load point")
    ap.add_argument("--debug-output", action="store_true",
                    help="make the renderer's line/point data after every step (option debug_output; "
                         "reports their counts; the timed steps include it)")
    ap.add_argument("--snapshot-dir", default=None,
                    help="after the timed steps, save the device world as a snapshot file in this folder and "
                         "load it into a fresh context (reports bytes and both rates; the file is removed)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 512 MiB device to device once and exit: the reference kernel of known traffic "
                         "(512 MiB read, 512 MiB written) whose FETCH_SIZE / WRITE_SIZE counters scale the "
                         "counters of the step kernels to HBM bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: only one GPU is supported in this build (asked for --gpus {args.gpus}); "
              "scenes are independent, so a multi-GPU run would shard whole scenes across GPUs", file=sys.stderr)
        sys.exit(2)

    from skyephyx_amd import roofline
    from skyephyx_amd.capi import Context

    t0 = time.perf_counter()
    world, desc = make_workload(args.workload)
    log(f"[bench] workload {args.workload}: {len(world.scenes)} scenes, {len(world.bodies)} bodies, "
        f"{len(world.tets)} tets, built in {time.perf_counter() - t0:.1f}s")
    ctx = Context(lanes=args.lanes, fem_lanes=args.fem_lanes)
    if args.no_graphs:
        ctx.set_option("step_graphs", 0)
    if args.fused:
        ctx.set_option("fused_scenes", 1)
    if args.debug_output:
        ctx.set_option("debug_output", 1)
    for opt in args.option:
        name, value = opt.split("=", 1)
        ctx.set_option(name, int(value))
    t0 = time.perf_counter()
    ctx.upload(world)
    ctx.synchronize()
    log(f"[bench] uploaded in {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    for i in range(0, args.settle, 10):
        ctx.step(min(10, args.settle - i))
    ctx.synchronize()
    log(f"[bench] settled {args.settle} steps in {time.perf_counter() - t0:.1f}s")
    for i in range(args.warmup):
        ctx.step(1)
    ctx.synchronize()
    log(f"[bench] warmup {args.warmup} done")
    # the timed steps: the product's defaults (step graphs, no per-launch events)
    ctx.kernel_stats("step_graph", reset=True)
    step_ms = []
    ctx.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        a = time.perf_counter()
        ctx.step(1)
        step_ms.append((time.perf_counter() - a) * 1e3)
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    graph_replays = ctx.kernel_stats("step_graph")[0]
    timings = ctx.timings()
    state = ctx.download()
    em, _ = ctx.status(len(state.scenes))
    # instrumented pass: HIP events on the stream each sweep launch runs on,
    # one pair per launch (the steps launch eagerly meanwhile)
    ks = None
    if not args.no_roofline and args.roofline_steps > 0:
        ctx.set_option("kernel_timing", 1)
        for k in KERNEL_CLASSES:
            ctx.kernel_stats(k, reset=True)
        for i in range(args.roofline_steps):
            ctx.step(1)
        ctx.synchronize()
        ks = {k: ctx.kernel_stats(k) for k in KERNEL_CLASSES}
        ctx.set_option("kernel_timing", 0)
        log(f"[bench] instrumented pass {args.roofline_steps} steps: " +
            ", ".join(f"{k} {n} launches {ms:.1f} ms" for k, (n, ms) in ks.items()))
    tune = ctx.tune_table()
    ctx_bytes = ctx.memory()
    debug = None
    if args.debug_output:
        lines, points = ctx.debug_output()
        debug = {"lines_per_step": int(len(lines)), "points_per_step": int(len(points)),
                 "bytes_per_step": int(lines.nbytes + points.nbytes)}
    snapshot = None
    if args.snapshot_dir:
        path = os.path.join(args.snapshot_dir, f"{args.workload}.avbdsnap")
        t0 = time.perf_counter()
        ctx.save_snapshot(path)
        t_save = time.perf_counter() - t0
    ctx.close()
    if args.snapshot_dir:
        size = os.path.getsize(path)
        c2 = Context(lanes=args.lanes, fem_lanes=args.fem_lanes)
        t0 = time.perf_counter()
        c2.load_snapshot(path)
        c2.synchronize()
        t_load = time.perf_counter() - t0
        c2.close()
        os.remove(path)
        snapshot = {"bytes": size, "save_s": round(t_save, 3), "save_GB_per_s": round(size / t_save / 1e9, 3),
                    "load_s": round(t_load, 3), "load_GB_per_s": round(size / t_load / 1e9, 3),
                    "note": "save = device download + file write; load = file read + upload (host records to "
                            "device arrays)"}
        log(f"[bench] snapshot {size / 1e9:.2f} GB: save {t_save:.2f} s, load {t_load:.2f} s")
    log(f"[bench] timed {args.steps} steps in {elapsed:.3f}s; emergency scenes: {int(em.sum())}")
    log("[bench] step ms: " + " ".join(f"{t:.1f}" for t in step_ms))

    S = len(state.scenes)
    scene_steps = S * args.steps / elapsed
    forces_bytes, fem_bytes = roofline.class_step_bytes(state)
    dual_bytes = roofline.step_bytes(state)[1]
    updates = roofline.body_updates_per_step(state)

    def kernel_roofline(cls):
        """Roofline object of one kernel class over the instrumented pass:
        algorithmic bytes per step (the state after the timed steps) x the
        pass's steps / the summed launch durations."""
        n, ms = ks[cls]
        step_b = {"primal_fem": fem_bytes, "primal_forces": forces_bytes, "dual": dual_bytes}[cls]
        ach = step_b * args.roofline_steps / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        return {
            "bound": "hbm",
            "kernel": KERNEL_NAMES[cls],
            "achieved": round(ach, 2),
            "peak": roofline.HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(ach / roofline.HBM_PEAK_GBPS, 5),
            "traffic": measured_traffic(args.workload, args.lanes, cls),
            "launches": n,
            "avg_launch_us": round(ms * 1e3 / max(n, 1), 3),
            "device_ms_per_step": round(ms / args.roofline_steps, 4),
            "algorithmic_bytes_per_launch": int(step_b * args.roofline_steps / max(n, 1)),
        }

    roof = None
    if ks is not None:
        # dominant kernel: the primal bucket class with the larger device time
        cls = "primal_fem" if ks["primal_fem"][1] >= ks["primal_forces"][1] else "primal_forces"
        other = "primal_forces" if cls == "primal_fem" else "primal_fem"
        roof = kernel_roofline(cls)
        roof["measured_over"] = (f"{args.roofline_steps} steps right after the timed ones, launched eagerly with "
                                 f"an event pair around every sweep launch on the stream it runs on")
        roof["other_bucket"] = kernel_roofline(other) if ks[other][0] else None
        roof["dual"] = kernel_roofline("dual")
    res = {
        "metric": METRIC,
        "value": round(scene_steps, 3),
        "unit": "scene-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded scene generators)",
        "config": {
            "workload": desc,
            "n_scenes": S,
            "n_bodies": int(len(state.bodies)),
            "n_dynamic": int(((state.bodies["flags"] & 1) == 0).sum()),
            "n_tets": int(len(state.tets)),
            "n_manifolds": int(len(state.manifolds)),
            "lanes_per_body": args.lanes or "autotuned",
            "lanes_per_fem_vertex": args.fem_lanes or "autotuned",
            "tuned_lanes_workgroup": {f"{'fem' if k[0] else 'forces'}/2^{k[1]}": list(v)
                                      for k, v in sorted(tune.items())},
            "colours_last_step": int(timings["n_colours"]),
            "options": {o.split("=", 1)[0]: int(o.split("=", 1)[1]) for o in args.option},
        },
        "roofline": roof,
        "p50_step_ms": round(float(np.percentile(step_ms, 50)), 4),
        "p99_step_ms": round(float(np.percentile(step_ms, 99)), 4),
        "body_vertex_updates_per_s": round(updates * args.steps / elapsed, 1),
        "phase_ms_last_step": {k: round(v, 4) for k, v in timings.items() if k.endswith("_ms")},
        "timed_state": f"after {args.settle} settling and {args.warmup} warm-up steps",
        "step_graph_replays": graph_replays,
    }
    if not args.no_plain and S > 1:
        log("[bench] plain per-scene comparison")
        plain = plain_gpu_rate(state)
        res["plain_scene_steps_per_s"] = round(plain, 3)
        res["batched_over_plain"] = round(scene_steps / plain, 3)
    res["device_bytes"] = ctx_bytes
    if not args.no_serve and S > 1:
        log("[bench] served form")
        points, sst, sat, stim = served(state, args.serve_load, args.serve_seconds)
        res["served"] = {
            "hbm_budget_bytes": int(200e9),
            "resident_bytes": int(sst["resident_bytes"]),
            "resident_scenes": int(sst["resident_scenes"]),
            "lanes": "forces pinned at 4 lanes; FEM-only particles tuned 4-32 lanes with canonical lane sums "
                     "(results independent of the mix)",
            "request": "U{1..8} steps of a uniformly drawn scene, Poisson arrivals",
            "saturation": sat,
            "points": points,
            "phase_ms_last_tick": {k: round(v, 4) for k, v in stim.items() if k.endswith("_ms")},
        }
        # headline point: the load nearest 0.8 of saturation
        main_pt = min(points, key=lambda p: abs(p["load_fraction"] - 0.8))
        res["served_headline_load"] = main_pt["load_fraction"]
        res["served_p50_latency_ms"] = main_pt["latency_ms_p50"]
        res["served_p99_latency_ms"] = main_pt["latency_ms_p99"]
        res["served_p99_tick_ms"] = main_pt["tick_ms_p99"]
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu_rate, cpu_desc, per, (ids, weights) = cpu_baseline(state)
        for sid, nb, nt, dt in per:
            log(f"[bench]   cpu scene {sid}: {nb} bodies {nt} tets {dt * 1e3:.2f} ms (median)")
        res["cpu_baseline"] = {
            "value": round(cpu_rate, 4),
            "unit": "scene-steps/s",
            "cores": 1,
            "kind": "port",
            "sample": cpu_desc,
            "host": cpu_model(),
        }
        res["gpu_over_cpu"] = round(scene_steps / cpu_rate, 2)
        if S > 1:
            log("[bench] CPU, independent scenes on all host cores")
            rate_all, threads, n_steps = cpu_all_cores(state, ids, weights)
            res["cpu_all_cores"] = {
                "value": round(rate_all, 4),
                "unit": "scene-steps/s",
                "cores": threads,
                "kind": "port",
                "sample": f"the same {len(ids)} scenes, one per thread over {threads} threads for 8 s "
                          f"({n_steps} steps); weighted median step time under contention",
            }
            res["gpu_over_cpu_all_cores"] = round(scene_steps / rate_all, 2)
    if not args.no_boxes20k and args.workload == "mix":
        log("[bench] 20k boxes: plain GPU step vs the CPU step on the landed world")
        res["boxes20k_plain_gpu_vs_cpu"] = boxes20k_plain_vs_cpu(args.no_cpu)
    if snapshot:
        res["snapshot"] = snapshot
    if debug:
        res["debug_output"] = debug
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
