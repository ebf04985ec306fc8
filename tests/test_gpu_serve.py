"""Served form on the GPU: a scene's result is the same alone and in the mix.

Scenes stepped by the request server (continuous batching: requests of
different lengths join and leave between ticks, scenes are evicted and
brought back under a small HBM budget) are compared with the same scene
stepped alone in its own context.  With the lane widths fixed, every body's
arithmetic is identical in both, so the comparison is bit-exact; with the
autotuned widths the accumulation order may differ and the single-step
parity bars apply.
"""
import numpy as np
import pytest

import oracle_py as O
from skyephyx_amd import scenes
from skyephyx_amd.capi import Context, footprint
from skyephyx_amd.serve import Server, poisson_run
from skyephyx_amd.world import World
from parity import POS_TOL, rot_err, pos_rel_err

pytestmark = pytest.mark.gpu


def _warm(world, steps):
    o = O.Oracle(world, fp64=False, order=O.NATURAL)
    o.step(steps)
    return o.save()


def small_mix():
    """Seven small scenes covering every force and energy kind; some start
    with live contacts (warmed on the oracle), so registration carries
    manifolds."""
    return [
        _warm(scenes.box_stack_3d(levels=3, drop=0.02), 10),
        scenes.joint_chain(5),
        _warm(scenes.spring_bridge(6), 5),
        scenes.mass_springs(4),
        scenes.fem_block(n=5, cell=0.25),
        scenes.cloth_sheet(n=6, spacing=0.25),
        _warm(scenes.pyramid_2d(), 20),
    ]


def alone(world, steps, lanes):
    """The scene stepped on its own, with the server's accumulation rules:
    lanes[0] lanes per body with forces, FEM-only particles with canonical
    lane sums at lanes[1] lanes (0: tuned; the width does not change them)."""
    with Context(lanes=lanes[0]) as ctx:
        ctx.set_option("fem_canonical_sums", 1)
        ctx.set_option("fem_lanes", lanes[1])
        ctx.upload(world)
        ctx.step(steps)
        return ctx.download()


def _manifolds_by_pair(w):
    return {(int(m["body_a"]), int(m["body_b"])): m for m in w.manifolds}


def assert_bit_equal(a: World, b: World, name):
    for f in ("pos", "rot", "vel", "ang_vel"):
        assert np.array_equal(a.bodies[f], b.bodies[f]), (name, f)
    assert np.array_equal(a.joints["lambda"], b.joints["lambda"]), name
    assert np.array_equal(a.springs["lambda"], b.springs["lambda"]), name
    ma, mb = _manifolds_by_pair(a), _manifolds_by_pair(b)
    assert set(ma) == set(mb), name
    for k in ma:
        n = int(ma[k]["n_contacts"])
        assert n == int(mb[k]["n_contacts"]), (name, k)
        for f in ("id", "lambda", "penalty", "stick"):
            assert np.array_equal(ma[k]["contacts"][f][:n], mb[k]["contacts"][f][:n]), (name, k, f)


def assert_close(a: World, b: World, name):
    assert pos_rel_err(a.bodies, b.bodies) <= POS_TOL, name
    assert rot_err(a.bodies, b.bodies) <= POS_TOL, name


def test_footprint_matches_allocation():
    ws = small_mix()
    mixed = World.concat(ws)
    for w in ws + [mixed]:
        with Context() as ctx:
            ctx.upload(w)
            assert ctx.memory() == footprint(w)


def _serve(worlds, steps, budget, lanes, stagger=True, second=True):
    """Register every scene, submit one request per scene (lengths `steps`),
    the later half after the first ticks, plus (second=True) another request
    for scene 0 that queues behind its first; tick until idle.  lanes None
    keeps the server's default.  Returns the final scenes and the stats."""
    with Server(budget_bytes=budget, lanes=lanes[0], fem_lanes=lanes[1]) as srv:
        ids = [srv.add_scenes(w)[0] for w in worlds]
        srv.preload()
        half = len(ids) // 2
        for i in ids[:half]:
            srv.submit(i, steps[i])
        if stagger:
            srv.tick()
            srv.tick()
        for i in ids[half:]:
            srv.submit(i, steps[i])
        if second:
            srv.submit(ids[0], 2)
        srv.run_until_idle()
        done = srv.poll()
        assert len(done) == len(ids) + (1 if second else 0)
        assert all(d["state"] == 2 and d["t_done"] >= d["t_start"] >= d["t_arrival"] for d in done)
        out = [srv.download_scene(i) for i in ids]
        return out, srv.stats()


STEPS = [3, 5, 2, 4, 3, 1, 6]


def test_alone_equals_mixed_bit_exact():
    ws = small_mix()
    served, st = _serve(ws, STEPS, 200e9, (4, 4))
    assert st["layout_rebuilds"] == 1
    for k, (w, g) in enumerate(zip(ws, served)):
        ref = alone(w, STEPS[k] + (2 if k == 0 else 0), (4, 4))
        assert_bit_equal(g, ref, f"scene {k}")


def test_eviction_round_trip_bit_exact():
    """A budget that holds the largest scene with little room to spare forces
    evictions and rebuilds between ticks; results are unchanged.  Not every
    scene fits at preload, so at least one residency change follows it."""
    ws = small_mix()
    budget = max(footprint(w) for w in ws) + 4096
    assert footprint(World.concat(ws)) > budget
    served, st = _serve(ws, STEPS, budget, (4, 4))
    print("eviction stats", st)
    assert st["layout_rebuilds"] >= 2, st
    assert st["resident_bytes"] <= budget
    for k, (w, g) in enumerate(zip(ws, served)):
        ref = alone(w, STEPS[k] + (2 if k == 0 else 0), (4, 4))
        assert_bit_equal(g, ref, f"scene {k}")


def test_alone_close_to_mixed_autotuned():
    """An autotuned width for bodies with forces follows the launch size, so
    their accumulation order differs between the mix and the scene alone:
    one step from the same state agrees to the single-step parity bar.
    (Over many steps, contact stick/slip decisions in stacks amplify
    rounding, which is why the server pins that width by default; FEM-only
    particles sum canonically at any width.)"""
    ws = small_mix()
    served, _ = _serve(ws, [1] * len(ws), 200e9, (0, 0), stagger=False, second=False)
    for k, (w, g) in enumerate(zip(ws, served)):
        ref = alone(w, 1, (0, 0))
        assert_close(g, ref, f"scene {k}")


def test_default_server_is_bit_exact_over_many_steps():
    """The server's defaults (4 lanes per body with forces, canonical FEM
    sums at a tuned width) keep even the contact-rich pyramid bit-identical
    to the scene alone over many steps, the scene alone running its FEM
    particles at another width (32 lanes)."""
    ws = small_mix()
    steps = [12] * len(ws)
    served, _ = _serve(ws, steps, 200e9, (None, None), second=False)
    for k, (w, g) in enumerate(zip(ws, served)):
        assert_bit_equal(g, alone(w, 12, (4, 32)), f"scene {k}")


def test_poisson_requests_all_complete():
    ws = small_mix()
    with Server(budget_bytes=200e9, max_slots=4) as srv:
        for w in ws:
            srv.add_scenes(w)
        srv.preload()
        res = poisson_run(srv, range(len(ws)), rate=200.0, duration=0.5, steps=(1, 4), seed=1)
        st = srv.stats()
    assert res["requests"] > 20
    assert st["queued"] == 0 and st["running"] == 0
    assert res["latency_ms_p99"] >= res["latency_ms_p50"] > 0
    assert res["scenes_per_tick_mean"] <= 4
