"""Edge cases of the device step: a world of static bodies only, an empty
scene between others, and a scene whose colouring needs more than 64
colours (the colouring's search beyond its 64-bit mask).  Each against the
oracle (colours bit-exact, positions within the parity bar)."""
import numpy as np
import pytest

import oracle_py as O
from parity import POS_TOL, colouring_ok, summary
from skyephyx_amd import scenes
from skyephyx_amd.capi import Context
from skyephyx_amd.world import MODEL_CUBE, World, WorldBuilder

pytestmark = pytest.mark.gpu


def _static_only():
    b = WorldBuilder()
    b.add_body(MODEL_CUBE, 1.0, 0.5, (0.0, -10.0, 0.0), (40.0, 1.0, 40.0), is_static=True)
    b.add_body(MODEL_CUBE, 1.0, 0.5, (0.0, -9.0, 0.0), (1.0, 1.0, 1.0), is_static=True)
    return b.build()


def _empty_scene():
    b = WorldBuilder()
    w = b.build()  # a scene with no bodies
    assert len(w.bodies) == 0
    return w


def _spring_clique(n=70):
    """n boxes on a 10 x k grid (3 m apart, inside the scene's 50 m bounds),
    every pair tied by a spring: the greedy colouring gives every box its own
    colour."""
    b = WorldBuilder()
    ids = [b.add_body(MODEL_CUBE, 1.0, 0.5, (3.0 * (i % 10) - 15.0, 3.0 * (i // 10), 0.0), (1.0, 1.0, 1.0))
           for i in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            b.add_spring(ids[i], (0.0, 0.0, 0.0), ids[j], (0.0, 0.0, 0.0), 10.0)
    return b.build()


def test_static_only_world():
    """Nothing moves; the broad phase still pairs touching static boxes (the
    reference's loop has no static test, solver.cpp:233-256), as the oracle."""
    world = _static_only()
    o = O.Oracle(world, fp64=False, order=O.COLOUR_MAJOR)
    o.step(3)
    ref = o.save()
    with Context(lanes=1, fem_lanes=1) as ctx:
        ctx.upload(world)
        ctx.step(3)
        g = ctx.download()
        np.testing.assert_array_equal(g.bodies["pos"], world.bodies["pos"])
        np.testing.assert_array_equal(g.bodies["rot"], world.bodies["rot"])
        res = summary(g, ref)
        assert res["pairs_equal"] and res["n_equal"] and res["ids_equal"], res
        assert len(g.manifolds) == len(ref.manifolds) == 1
        assert np.all(ctx.colours(len(g.bodies)) == -1)


def test_empty_scene_between_scenes():
    a, c = scenes.pyramid_2d(rows=4), scenes.joint_chain(3)
    world = World.concat([a, _empty_scene(), c])
    assert len(world.scenes) == 3
    with Context(lanes=1, fem_lanes=1) as ctx, Context(lanes=1, fem_lanes=1) as alone:
        ctx.upload(world)
        alone.upload(c)
        ctx.step(4)
        alone.step(4)
        g = ctx.download()
        ctx.step_scenes([1], 2)  # only the empty scene: nothing moves
        h = ctx.download()
        solo = alone.download()
    np.testing.assert_array_equal(g.bodies["pos"], h.bodies["pos"])
    np.testing.assert_array_equal(g.extract_scene(2).bodies["pos"], solo.bodies["pos"])


def test_more_than_64_colours():
    world = _spring_clique(70)
    o = O.Oracle(world, fp64=False, order=O.COLOUR_MAJOR)
    with Context(lanes=1, fem_lanes=1) as ctx:
        ctx.upload(world)
        for k in range(3):
            ctx.step(1)
            o.step(1)
            g, ref = ctx.download(), o.save()
            gcol = ctx.colours(len(g.bodies))
            ocol, nc = o.colours(len(ref.bodies))
            assert nc == 70 and gcol.max() == 69
            np.testing.assert_array_equal(gcol, ocol)
            assert colouring_ok(g, gcol)
            res = summary(g, ref)
            assert res["pos_rel"] <= POS_TOL and res["rot"] <= POS_TOL, (k, res)


def test_element_records_in_any_order():
    """Tets and triangles handed over out of creation order: the upload
    orders each vertex's energies by serial and gathers the element stream
    on the device from the records, so the step is bit-identical to the one
    of the records in order."""
    world = World.concat([scenes.fem_block(n=3, cell=0.5), scenes.cloth_sheet(n=4, spacing=0.5)])
    shuffled = world.copy()
    rng = np.random.default_rng(5)
    shuffled.tets = world.tets[rng.permutation(len(world.tets))]
    shuffled.tris = world.tris[rng.permutation(len(world.tris))]
    out = []
    for w in (world, shuffled):
        with Context(lanes=1, fem_lanes=1) as ctx:
            ctx.upload(w)
            ctx.step(3)
            out.append(ctx.download())
    assert len(world.tets) > 0 and len(world.tris) > 0
    np.testing.assert_array_equal(out[0].bodies["pos"], out[1].bodies["pos"])
    np.testing.assert_array_equal(out[0].bodies["rot"], out[1].bodies["rot"])


def _two_fem_scenes_with_a_crossing_tet():
    w = World.concat([scenes.fem_block(n=2, cell=0.5), scenes.fem_block(n=2, cell=0.5, origin=(3.0, 0.0, 0.0))])
    first = w.bodies["scene"] == 0
    other = int(np.nonzero(~first)[0][0])
    t = int(np.nonzero(np.all(np.isin(w.tets["v"], np.nonzero(first)[0]), axis=1))[0][0])
    w.tets["v"][t][3] = other  # one vertex in the second scene
    return w


def test_cross_scene_element_is_refused_before_any_launch():
    """An element whose vertices lie in two scenes is refused by the upload
    (AVBD_E_ARG) before anything is freed or launched: the world uploaded
    before keeps stepping as a context that never saw the bad upload."""
    from skyephyx_amd.capi import AvbdError

    good = scenes.fem_block(n=3, cell=0.5)
    with Context(lanes=1, fem_lanes=1) as a, Context(lanes=1, fem_lanes=1) as b:
        a.upload(good)
        b.upload(good)
        a.step(2)
        with pytest.raises(AvbdError, match="two scenes"):
            a.upload(_two_fem_scenes_with_a_crossing_tet())
        cloth = World.concat([scenes.cloth_sheet(n=3, spacing=0.5), scenes.cloth_sheet(n=3, spacing=0.5)])
        cloth.tris["v"][0][2] = int(np.nonzero(cloth.bodies["scene"] == 1)[0][0])
        with pytest.raises(AvbdError, match="two scenes"):
            a.upload(cloth)
        a.step(2)
        b.step(4)
        ga, gb = a.download(), b.download()
        assert ga.bodies.tobytes() == gb.bodies.tobytes()


def test_spare_slots_do_not_widen_the_index_range():
    """With spare body slots, an index at or beyond the caller's body count
    (which would name a spare or shifted body after the expansion) is
    refused, as without spare slots."""
    from skyephyx_amd.capi import AvbdError

    w = World.concat([scenes.joint_chain(3), scenes.joint_chain(3)])
    w.joints["body_b"][0] = len(w.bodies)  # one past the last body
    with Context() as ctx:
        ctx.set_option("spare_bodies", 4)
        with pytest.raises(AvbdError, match="joint body index"):
            ctx.upload(w)
        w.joints["body_b"][0] = -2
        with pytest.raises(AvbdError, match="joint body index"):
            ctx.upload(w)
