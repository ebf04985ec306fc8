"""Out-of-bounds removal on the GPU against the oracle (solver.cpp:520-534).

Bodies that leave the world bounds during a step are removed before its
post-stabilisation, each replaced by the scene's last body (swap with last),
together with their forces and energies.  The GPU renumbers its device state
the same way; over three steps the surviving bodies, their order, the forces,
energies and contacts must match the oracle in colour-major order.
"""
import numpy as np
import pytest

import oracle_py as O
from skyephyx_amd import scenes
from skyephyx_amd.capi import Context
from skyephyx_amd.world import World
from parity import POS_TOL, summary

pytestmark = pytest.mark.gpu


def _counts(w):
    return (len(w.bodies), len(w.joints), len(w.springs), len(w.tets), len(w.tris), int(w.scenes["n_removed"].sum()))


@pytest.mark.parametrize("lanes", [(1, 1), (0, 0)], ids=["seq", "tuned"])
def test_removal_matches_oracle(lanes):
    world = scenes.out_of_bounds()
    o = O.Oracle(world, fp64=False, order=O.COLOUR_MAJOR)
    with Context(lanes=lanes[0], fem_lanes=lanes[1]) as ctx:
        ctx.upload(world)
        for step in range(3):
            ctx.step(1)
            o.step(1)
            g, ref = ctx.download(), o.save()
            assert _counts(g) == _counts(ref), (step, _counts(g), _counts(ref))
            # the survivors, in the same order
            np.testing.assert_array_equal(g.bodies["flags"], ref.bodies["flags"])
            np.testing.assert_array_equal(g.bodies["mass"], ref.bodies["mass"])
            res = summary(g, ref)
            assert res["pairs_equal"] and res["n_equal"] and res["ids_equal"], (step, res)
            assert res["pos_rel"] <= POS_TOL and res["rot"] <= POS_TOL, (step, res)
            np.testing.assert_array_equal(np.sort(g.tets["v"], axis=1), np.sort(ref.tets["v"], axis=1))


def test_removal_in_one_scene_leaves_the_others_alone():
    """Two copies of the scene around a resting pyramid: removal renumbers
    every later scene's bodies; the pyramid's state is the same as alone."""
    pyr = scenes.pyramid_2d(rows=4)
    world = World.concat([scenes.out_of_bounds(), pyr, scenes.out_of_bounds()])
    with Context(lanes=4, fem_lanes=4) as ctx:
        ctx.upload(world)
        ctx.step(3)
        both = ctx.download()
    with Context(lanes=4, fem_lanes=4) as ctx:
        ctx.upload(pyr)
        ctx.step(3)
        alone = ctx.download()
    mid = both.extract_scene(1)
    np.testing.assert_array_equal(mid.bodies["pos"], alone.bodies["pos"])
    np.testing.assert_array_equal(mid.bodies["rot"], alone.bodies["rot"])
    assert int(both.scenes["n_removed"][0]) == int(both.scenes["n_removed"][2]) == 9
    # both removing scenes end as the scene does alone: same order, poses,
    # surviving tets (renumbered) and forces
    with Context(lanes=4, fem_lanes=4) as ctx:
        ctx.upload(scenes.out_of_bounds())
        ctx.step(3)
        oob = ctx.download()
    for k in (0, 2):
        sc = both.extract_scene(k)
        for f in ("pos", "rot", "vel", "ang_vel"):
            np.testing.assert_array_equal(sc.bodies[f], oob.bodies[f])
        np.testing.assert_array_equal(sc.tets["v"], oob.tets["v"])
        np.testing.assert_array_equal(sc.springs["body_b"], oob.springs["body_b"])
        np.testing.assert_array_equal(sc.manifolds["body_a"], oob.manifolds["body_a"])
        np.testing.assert_array_equal(sc.manifolds["body_b"], oob.manifolds["body_b"])


def _emulate(n, ids):
    """Body order after RemoveBody of ids (indices before the call), and the
    index each had when it was removed."""
    order = list(range(n))
    at = []
    for i in ids:
        k = order.index(i)
        at.append(k)
        order[k] = order[-1]
        order.pop()
    return order, at


def _warm(world, steps):
    o = O.Oracle(world, fp64=False, order=O.NATURAL)
    o.step(steps)
    return o.save()


EXPLICIT = {
    # boxes with contacts, a joint chain and springs
    "lattice_mix": (lambda: _warm(scenes.box_lattice(shape=(4, 3, 3), spacing=3.0, n_chain=12, n_springs=8,
                                                     bounds=50.0), 20), lambda n: [5, 0, n - 1, 40]),
    # FEM vertices: their tets go with them
    "fem_block": (lambda: scenes.fem_block(n=5, cell=0.25), lambda n: [62, 0, 31]),
    "pyramid": (lambda: _warm(scenes.pyramid_2d(), 30), lambda n: [n - 1, 10, 1]),
}


@pytest.mark.parametrize("case", sorted(EXPLICIT))
def test_explicit_removal_matches_oracle(case):
    """avbd_remove_bodies on the device-resident world equals the oracle's
    RemoveBody sequence, and the next step keeps the parity bars."""
    make, pick = EXPLICIT[case]
    world = make()
    ids = pick(len(world.bodies))
    o = O.Oracle(world, fp64=False, order=O.COLOUR_MAJOR)
    with Context(lanes=1, fem_lanes=1) as ctx:
        ctx.upload(world)
        ctx.remove_bodies(ids)
        o.remove_bodies(ids)
        g, ref = ctx.download(), o.save()
        order, at = _emulate(len(world.bodies), ids)
        assert ctx.last_removed() == [(0, k) for k in at]
        assert _counts(g) == _counts(ref), (_counts(g), _counts(ref))
        for f in ("pos", "rot", "vel", "ang_vel", "flags", "mass"):
            np.testing.assert_array_equal(g.bodies[f], ref.bodies[f], err_msg=f)
        np.testing.assert_array_equal(g.bodies["pos"], world.bodies["pos"][order])
        np.testing.assert_array_equal(np.sort(g.tets["v"], axis=1), np.sort(ref.tets["v"], axis=1))
        res = summary(g, ref)
        assert res["pairs_equal"] and res["n_equal"] and res["ids_equal"], res
        ctx.step(1)
        o.step(1)
        g, ref = ctx.download(), o.save()
        assert _counts(g) == _counts(ref)
        res = summary(g, ref)
        assert res["pairs_equal"] and res["n_equal"] and res["ids_equal"], res
        assert res["pos_rel"] <= POS_TOL and res["rot"] <= POS_TOL, res


def test_explicit_removal_across_scenes():
    """Removals in two scenes of one world, interleaved in the list: each
    scene ends as its own oracle does; the scene between is untouched."""
    a = _warm(scenes.pyramid_2d(rows=5), 20)
    mid = scenes.joint_chain(4)
    b = scenes.fem_block(n=4, cell=0.25)
    world = World.concat([a, mid, b])
    na, nm = len(a.bodies), len(mid.bodies)
    ids = [na + nm + 5, 3, na + nm + 0, 7]  # b:5, a:3, b:0, a:7
    with Context(lanes=1, fem_lanes=1) as ctx:
        ctx.upload(world)
        ctx.remove_bodies(ids)
        g = ctx.download()
        assert [s for s, _ in ctx.last_removed()] == [0, 0, 2, 2]
    for scene, sub, local in ((0, a, [3, 7]), (2, b, [5, 0])):
        o = O.Oracle(sub, fp64=False, order=O.COLOUR_MAJOR)
        o.remove_bodies(local)
        ref = o.save()
        got = g.extract_scene(scene)
        np.testing.assert_array_equal(got.bodies["pos"], ref.bodies["pos"])
        assert len(got.manifolds) == len(ref.manifolds) and len(got.tets) == len(ref.tets)
        np.testing.assert_array_equal(np.sort(got.tets["v"], axis=1), np.sort(ref.tets["v"], axis=1))
    np.testing.assert_array_equal(g.extract_scene(1).bodies["pos"], mid.bodies["pos"])


def test_add_bodies_into_freed_slots():
    """avbd_add_bodies appends bodies to a scene's body order in the slots a
    removal freed: the world then equals the oracle's after RemoveBody and
    AddBody (a body appended to the records), and steps within the parity
    bars.  A scene with no freed slot left refuses with AVBD_E_CAPACITY."""
    world = _warm(scenes.pyramid_2d(), 30)
    ids = [12, 30]
    o = O.Oracle(world, fp64=False, order=O.COLOUR_MAJOR)
    o.remove_bodies(ids)
    removed = o.save()
    new = world.bodies[[5, 6]].copy()  # two unit boxes, dropped above the pile
    new["pos"] = [(0.3, 14.0, 0.0), (-0.4, 16.0, 0.0)]
    for f in ("vel", "ang_vel", "prev_vel", "prev_ang_vel"):
        new[f] = 0.0
    expect = World(removed.scenes, np.concatenate([removed.bodies, new]), removed.joints, removed.springs,
                   removed.manifolds, removed.tets, removed.tris)
    with Context(lanes=1, fem_lanes=1) as ctx:
        ctx.upload(world)
        ctx.remove_bodies(ids)
        ctx.add_bodies(new)
        g = ctx.download()
        assert len(g.bodies) == len(expect.bodies)
        assert g.bodies.tobytes() == expect.bodies.tobytes()
        with pytest.raises(Exception) as e:
            ctx.add_bodies(new[:1])  # both freed slots are taken
        assert getattr(e.value, "code", None) == -3
        ref = O.Oracle(expect, fp64=False, order=O.COLOUR_MAJOR)
        for _ in range(3):
            ctx.step(1)
            ref.step(1)
            g, r = ctx.download(), ref.save()
            res = summary(g, r)
            assert res["pairs_equal"] and res["n_equal"] and res["ids_equal"], res
            assert res["pos_rel"] <= POS_TOL and res["rot"] <= POS_TOL, res


def test_spare_slots_change_nothing_and_take_added_bodies():
    """Option spare_bodies: free slots after every scene's bodies.  They are
    invisible (downloads, counts, removals) and the steps are bit-identical
    to a context without them; avbd_add_bodies fills them without any
    removal, as the oracle's world with the record appended."""
    world = World.concat([_warm(scenes.pyramid_2d(rows=6), 20), scenes.fem_block(n=3, cell=0.5),
                          scenes.joint_chain(4)])
    with Context(lanes=1, fem_lanes=1) as sp, Context(lanes=1, fem_lanes=1) as plain:
        sp.set_option("spare_bodies", 3)
        sp.upload(world)
        plain.upload(world)
        g = sp.download()
        for f in ("bodies", "joints", "tets", "manifolds"):
            assert getattr(g, f).tobytes() == getattr(world, f).tobytes(), f
        for _ in range(6):
            sp.step(1)
            plain.step(1)
        a, b = sp.download(), plain.download()
        for f in ("bodies", "joints", "springs", "tets"):
            assert getattr(a, f).tobytes() == getattr(b, f).tobytes(), f
        assert len(a.manifolds) == len(b.manifolds)
        # three boxes into scene 0's spare slots; a fourth has no slot
        new = a.bodies[[3, 4, 5]].copy()
        new["pos"] = [(0.0, 12.0, 0.0), (0.5, 14.0, 0.0), (-0.5, 16.0, 0.0)]
        for f in ("vel", "ang_vel", "prev_vel", "prev_ang_vel"):
            new[f] = 0.0
        sp.add_bodies(new)
        with pytest.raises(Exception) as e:
            sp.add_bodies(new[:1])
        assert getattr(e.value, "code", None) == -3
        got = sp.download()
        n0 = int((a.bodies["scene"] == 0).sum())
        want = np.concatenate([a.bodies[:n0], new, a.bodies[n0:]])
        assert got.bodies.tobytes() == want.tobytes()
        scene0 = got.extract_scene(0)
        ref = O.Oracle(scene0, fp64=False, order=O.COLOUR_MAJOR)
        sp.step(1)
        ref.step(1)
        res = summary(sp.download().extract_scene(0), ref.save())
        assert res["pairs_equal"] and res["n_equal"] and res["ids_equal"], res
        assert res["pos_rel"] <= POS_TOL and res["rot"] <= POS_TOL, res


@pytest.mark.parametrize("lanes", [(1, 1), (0, 0)], ids=["seq", "tuned"])
def test_drag_matches_oracle(lanes):
    """Mesh::isDragged / addedDragVelocity (solver.cpp:506-510) set on the
    resident world: a dragged box of a settled pile and one of a joint chain
    get their added velocity once in the next velocity update, the flag stays
    set, an undragged body ignores its added velocity.  Each step against the
    oracle given the same drag state; also after a removal renumbered the
    bodies (the drag state moves with its body)."""
    o = O.Oracle(World.concat([scenes.box_stack_3d(levels=3, drop=0.02)]), fp64=False, order=O.NATURAL)
    o.step(15)
    state = o.save()
    n = len(state.bodies)
    with Context(lanes=lanes[0], fem_lanes=lanes[1]) as ctx:
        ctx.upload(state)
        ref = O.Oracle(state, fp64=False, order=O.COLOUR_MAJOR)
        drag = [(n - 1, (1.5, 2.0, 0.0), 1), (n - 3, (0.0, 0.0, -3.0), 1), (2, (9.0, 0.0, 0.0), 0)]
        for i, v, d in drag:
            ctx.set_drag([i], v, d)
            ref.set_drag([i], v, d)
        for step in range(3):
            ctx.step(1)
            ref.step(1)
            g, r = ctx.download(), ref.save()
            np.testing.assert_allclose(g.bodies["vel"], r.bodies["vel"], rtol=0, atol=2e-3)
            np.testing.assert_allclose(g.bodies["pos"], r.bodies["pos"], rtol=0, atol=1e-4)
            assert len(g.manifolds) == len(r.manifolds), step
        # the dragged box left with about its drag added
        assert g.bodies["vel"][n - 1][0] > 0.5
        # a removal renumbers: body n-1 swaps into slot 0, keeps its (spent) drag flag
        ctx.remove_bodies([0])
        ref.remove_bodies([0])
        ctx.set_drag([0], (0.0, 4.0, 0.0), 1)
        ref.set_drag([0], (0.0, 4.0, 0.0), 1)
        ctx.step(1)
        ref.step(1)
        g, r = ctx.download(), ref.save()
        np.testing.assert_allclose(g.bodies["vel"], r.bodies["vel"], rtol=0, atol=2e-3)
        np.testing.assert_allclose(g.bodies["pos"], r.bodies["pos"], rtol=0, atol=1e-4)
