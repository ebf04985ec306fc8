"""Colouring updates on the GPU.

After its first colouring a scene whose constraint graph gains or loses a few
edges (contacts appearing and vanishing, a joint or spring breaking) updates
its last colouring from the bodies the change can reach instead of colouring
from scratch.  The greedy colouring is unique, so over many steps the updated
colouring must equal the one from scratch (option colour_updates 0): same
colours, and therefore bit-identical states.  Each step's colours are also
checked against the greedy rule on the step's own constraint graph.
"""
import numpy as np
import pytest

import oracle_py as O
from parity import colouring_ok
from skyephyx_amd import scenes, scenes_fast
from skyephyx_amd.capi import Context
from skyephyx_amd.world import World

pytestmark = pytest.mark.gpu


def _warm(world, steps):
    o = O.Oracle(world, fp64=False, order=O.NATURAL)
    o.step(steps)
    return o.save()


CASES = {
    # boxes falling onto each other and the ground: contacts appear every step
    "box_pile_600": (lambda: scenes_fast.box_pile(600, np.random.default_rng(11)), 24),
    # boxes, a ball-joint chain and springs
    "lattice_mix": (lambda: scenes.box_lattice(shape=(4, 3, 3), spacing=3.0, n_chain=12, n_springs=8, bounds=50.0), 40),
    "pyramid_settling": (lambda: scenes.pyramid_2d(), 30),
    # springs that break (the graph loses edges through the fracture path)
    "spring_bridge": (lambda: scenes.spring_bridge(6), 30),
    "fem_on_ground": (lambda: scenes.fem_block(n=6, cell=0.25, pin_top=False, ground=True,
                                               origin=(-0.6, -9.45, -0.6)), 12),
    # several scenes in one world
    "mixed_world": (lambda: World.concat([
        _warm(scenes.pyramid_2d(rows=6), 10),
        scenes_fast.box_pile(300, np.random.default_rng(3)),
        scenes.joint_chain(5),
        scenes.spring_bridge(4),
    ]), 20),
}


def _same_state(a: World, b: World, where):
    assert len(a.bodies) == len(b.bodies), where
    for f in ("pos", "rot", "vel", "ang_vel"):
        assert np.array_equal(a.bodies[f], b.bodies[f]), (where, f)
    assert len(a.manifolds) == len(b.manifolds), where


@pytest.mark.parametrize("case", list(CASES))
def test_colour_updates_equal_colouring_from_scratch(case):
    make, steps = CASES[case]
    world = make()
    n = len(world.bodies)
    with Context(lanes=4, fem_lanes=4) as upd, Context(lanes=4, fem_lanes=4) as full:
        upd.set_option("colour_updates", 1)
        full.set_option("colour_updates", 0)
        upd.upload(world)
        full.upload(world)
        for k in range(steps):
            upd.step(1)
            full.step(1)
            cu = upd.colours(n)
            cf = full.colours(n)
            assert np.array_equal(cu, cf), (case, k, int(np.sum(cu != cf)))
            if k % 6 == 5 or k == steps - 1:
                a = upd.download()
                _same_state(a, full.download(), (case, k))
                assert colouring_ok(a, cu[:len(a.bodies)]), (case, k)


def test_colour_updates_over_scene_subsets():
    """Scenes stepped in changing subsets (the served form): a scene's edge
    changes are only recorded while it runs."""
    world = World.concat([scenes_fast.box_pile(200, np.random.default_rng(5)), scenes.pyramid_2d(rows=5),
                          scenes_fast.box_pile(150, np.random.default_rng(6)), scenes.spring_bridge(4)])
    n = len(world.bodies)
    subsets = [[0, 1, 2, 3], [0, 2], [1, 3], [0], [2, 3], [0, 1, 2, 3], [1], [0, 2]] * 3
    with Context(lanes=4, fem_lanes=4) as upd, Context(lanes=4, fem_lanes=4) as full:
        upd.set_option("colour_updates", 1)
        full.set_option("colour_updates", 0)
        upd.upload(world)
        full.upload(world)
        for k, ids in enumerate(subsets):
            upd.step_scenes(ids, 1)
            full.step_scenes(ids, 1)
            assert np.array_equal(upd.colours(n), full.colours(n)), k
        _same_state(upd.download(), full.download(), "end")


def test_colour_updates_large_scene():
    """A scene above the LDS size limit (colours kept in global memory)."""
    world = scenes_fast.box_pile(40000, np.random.default_rng(7))
    n = len(world.bodies)
    with Context(lanes=4, fem_lanes=4) as upd, Context(lanes=4, fem_lanes=4) as full:
        upd.set_option("colour_updates", 1)
        full.set_option("colour_updates", 0)
        upd.upload(world)
        full.upload(world)
        for k in range(8):
            upd.step(1)
            full.step(1)
            assert np.array_equal(upd.colours(n), full.colours(n)), k
        _same_state(upd.download(), full.download(), "end")
