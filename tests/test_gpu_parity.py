"""GPU-vs-oracle parity of one AVBD step from an identical state.

The HIP path runs through the C ABI; the oracle runs the same step in
colour-major visit order (bodies of one colour share no force, so this is
exactly what the per-colour parallel sweep computes).  Bars: positions and
orientations within 1e-4 relative, lambda within 1e-3 relative; contact
pairs, counts, feature IDs and colours bit-exact.  The FP64 oracle error of
the same step is reported beside it, and the distance from the FP32 oracle
in the reference's natural visit order as a diagnostic.
"""
import json
import os

import numpy as np
import pytest

import oracle_py as O
from skyephyx_amd import scenes, scenes_fast
from skyephyx_amd.capi import Context
from skyephyx_amd.world import World
from parity import assert_parity, one_step_report

pytestmark = pytest.mark.gpu


def warm_state(world, steps):
    """Advance a world with the FP32 oracle (natural order) to get a state
    with live contacts and warm-started multipliers."""
    if steps == 0:
        return world
    o = O.Oracle(world, fp64=False, order=O.NATURAL)
    o.step(steps)
    return o.save()


def _record(name, res):
    out = os.environ.get("PARITY_OUT")
    print(name, json.dumps(res, default=str))
    if out:
        with open(out, "a") as f:
            f.write(json.dumps({"case": name, **res}, default=str) + "\n")


CASES = {
    "pyramid_c1_step0": (lambda: scenes.pyramid_2d(), 0),
    "pyramid_c1_step40": (lambda: scenes.pyramid_2d(), 40),
    "box_stack_3d": (lambda: scenes.box_stack_3d(levels=3, drop=0.02), 30),
    "joint_chain": (lambda: scenes.joint_chain(5), 10),
    "spring_bridge": (lambda: scenes.spring_bridge(6), 20),
    "mass_springs": (lambda: scenes.mass_springs(4), 15),
    "fem_block": (lambda: scenes.fem_block(n=6, cell=0.25), 5),
    "cloth": (lambda: scenes.cloth_sheet(n=8, spacing=0.25), 5),
    "lattice_mix": (lambda: scenes.box_lattice(shape=(4, 3, 3), spacing=3.0, n_chain=12, n_springs=8, bounds=50.0), 50),
    # several broad-phase tiles of 256 bodies: contacts between tiles and
    # against the ground in the first tile
    "box_pile_600": (lambda: scenes_fast.box_pile(600, np.random.default_rng(11)), 30),
    "box_pile_600_step0": (lambda: scenes_fast.box_pile(600, np.random.default_rng(11)), 0),
    "fem_on_ground": (lambda: scenes.fem_block(n=9, cell=0.25, pin_top=False, ground=True,
                                               origin=(-1.0, -9.45, -1.0)), 5),
    "fem_on_ground_step0": (lambda: scenes.fem_block(n=9, cell=0.25, pin_top=False, ground=True,
                                                     origin=(-1.0, -9.45, -1.0)), 0),
}


@pytest.mark.parametrize("lanes", [(1, 1), (8, 16), (0, 0)], ids=["seq", "coop", "tuned"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_single_step_parity(case, lanes):
    """lanes = (lanes per body with forces, lanes per FEM-only body); (1, 1)
    accumulates in the reference's sequential order and is held to the
    lambda bar per entry."""
    make, warm = CASES[case]
    state = warm_state(make(), warm)
    ctx = Context(lanes=lanes[0], fem_lanes=lanes[1])
    ctx.upload(state)
    ctx.step(1)
    g = ctx.download()
    gcol = ctx.colours(len(g.bodies))
    res = one_step_report(state, g, gcol)
    _record(f"{case}/lanes{lanes[0]}-{lanes[1]}", res)
    assert_parity(res, strict_lambda=lanes == (1, 1))


def _split(world, gcol):
    """Per scene: (one-scene world, its bodies' colours)."""
    out = []
    for s in range(len(world.scenes)):
        idx = world.scene_bodies(s)
        out.append((world.extract_scene(s), gcol[idx] if len(idx) else gcol[:0]))
    return out


# scenes of one world with their own time step and iteration count (the
# batched mix gives every scene dt in {1/60, 1/120} and 8-16 iterations)
MIXED_PARAMS = [
    ("pyramid", lambda **p: scenes.pyramid_2d(**p), 30, dict(dt=1.0 / 60.0, iterations=8)),
    ("fem_block", lambda **p: scenes.fem_block(n=5, cell=0.25, **p), 5, dict(dt=1.0 / 120.0, iterations=16)),
    ("joint_chain", lambda **p: scenes.joint_chain(5, **p), 10, dict(dt=1.0 / 120.0, iterations=11)),
    ("spring_bridge", lambda **p: scenes.spring_bridge(6, **p), 20, dict(dt=1.0 / 60.0, iterations=13)),
    ("cloth", lambda **p: scenes.cloth_sheet(n=6, spacing=0.25, **p), 5, dict(dt=1.0 / 60.0, iterations=9)),
    ("box_pile", lambda **p: scenes_fast.box_pile(120, np.random.default_rng(4), **p), 25,
     dict(dt=1.0 / 120.0, iterations=16)),
    ("fem_on_ground", lambda **p: scenes.fem_block(n=4, cell=0.25, pin_top=False, ground=True,
                                                   origin=(-0.5, -9.45, -0.5), **p), 5,
     dict(dt=1.0 / 60.0, iterations=12)),
]


@pytest.mark.parametrize("lanes", [(1, 1), (0, 0)], ids=["seq", "tuned"])
def test_scenes_with_their_own_dt_and_iterations(lanes):
    """One world of seven scenes whose dt and iteration counts differ (the
    sweeps run the longest scene's iterations; a scene stops at its own
    count): one step of the world against each scene stepped alone by the
    oracle with its own dt and iterations."""
    states = []
    for name, make, warm, params in MIXED_PARAMS:
        w = make(**{k: (np.float32(v) if k == "dt" else v) for k, v in params.items()})
        assert w.scenes["iterations"][0] == params["iterations"], name
        states.append(warm_state(w, warm))
    world = World.concat(states)
    with Context(lanes=lanes[0], fem_lanes=lanes[1]) as ctx:
        ctx.upload(world)
        ctx.step(1)
        g = ctx.download()
        gcol = ctx.colours(len(g.bodies))
    for (name, _, _, params), state, (gs, gc) in zip(MIXED_PARAMS, states, _split(g, gcol)):
        # colours are numbered per scene on the device as in the oracle
        res = one_step_report(state, gs, gc, natural=False)
        _record(f"mixed_params/{name}/lanes{lanes[0]}-{lanes[1]}", res)
        assert_parity(res, strict_lambda=lanes == (1, 1))


def test_explosion_stop_against_the_oracle():
    """CheckExplosion (solver.cpp:182-196, 528-529): a body leaving at more
    than 500 m/s stops its scene after the velocity update, before the
    bounds removal and post-stabilisation, and later steps leave it as it
    is; another scene in the same world goes on.  Each scene against the
    oracle stepping it alone."""
    pile = warm_state(scenes.box_stack_3d(levels=3, drop=0.02), 10)
    boom = pile.copy()
    dyn = np.nonzero((boom.bodies["flags"] & 1) == 0)[0]
    boom.bodies["vel"][dyn[-1]] = (700.0, 0.0, 0.0)
    boom.bodies["pos"][dyn[0]] = (60.0, 0.0, 0.0)  # out of bounds, but the stop comes first
    other = warm_state(scenes.joint_chain(5), 10)
    world = World.concat([boom, other])
    with Context(lanes=1, fem_lanes=1) as ctx:
        ctx.upload(world)
        ctx.step(1)
        g1 = ctx.download()
        gcol = ctx.colours(len(g1.bodies))
        ctx.step(3)
        g4 = ctx.download()
    o = O.Oracle(boom, fp64=False, order=O.COLOUR_MAJOR)
    o.step(1)
    ref = o.save()
    assert ref.scenes["emergency_stop"][0] == 1 and ref.scenes["n_removed"][0] == 0
    parts = _split(g1, gcol)
    gs, gc = parts[0]
    assert gs.scenes["emergency_stop"][0] == 1
    assert gs.scenes["n_removed"][0] == 0 and len(gs.bodies) == len(ref.bodies)
    res = one_step_report(boom, gs, gc, natural=False)
    _record("explosion/stopped_scene", res)
    assert_parity(res, strict_lambda=True)
    res = one_step_report(other, parts[1][0], parts[1][1], natural=False)
    _record("explosion/other_scene", res)
    assert_parity(res, strict_lambda=True)
    # the stopped scene no longer moves; the other one does
    s0, s1 = g4.extract_scene(0), g4.extract_scene(1)
    assert s0.bodies.tobytes() == gs.bodies.tobytes()
    assert s1.bodies["pos"].tobytes() != parts[1][0].bodies["pos"].tobytes()
