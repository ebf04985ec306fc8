"""Step graphs on the GPU: phases 4-7 of a step replayed from a captured HIP
graph give exactly the state of launching them one by one.

A step's launch shape (colours, iterations, lane widths, grids) is captured
the second time it occurs and replayed after that; the kernels read the
bucket bounds and the manifold count from device memory, so a replay follows
the step it runs in.  Every step is compared bit for bit with a context that
never uses graphs, with the download (a synchronising read) between replays.
"""
import numpy as np
import pytest

import oracle_py as O
from skyephyx_amd import scenes
from skyephyx_amd.capi import Context
from skyephyx_amd.world import World

pytestmark = pytest.mark.gpu


def _warm(world, steps):
    o = O.Oracle(world, fp64=False, order=O.NATURAL)
    o.step(steps)
    return o.save()


def _equal(a: World, b: World, where):
    assert len(a.bodies) == len(b.bodies), where
    for f in ("pos", "rot", "vel", "ang_vel"):
        assert np.array_equal(a.bodies[f], b.bodies[f]), (where, f)
    assert len(a.manifolds) == len(b.manifolds), where
    assert np.array_equal(a.manifolds["body_a"], b.manifolds["body_a"]), where
    for f in ("lambda", "penalty"):
        assert np.array_equal(a.manifolds["contacts"][f], b.manifolds["contacts"][f]), (where, f)
    assert np.array_equal(a.joints["lambda"], b.joints["lambda"]), where
    assert np.array_equal(a.springs["lambda"], b.springs["lambda"]), where


CASES = {
    "pyramid": lambda: scenes.pyramid_2d(),
    "pyramid_settled": lambda: _warm(scenes.pyramid_2d(), 40),
    "spring_bridge_fracture": lambda: scenes.spring_bridge(6),
    "joint_chain": lambda: scenes.joint_chain(8),
    "fem_and_boxes": lambda: World.concat([
        scenes.fem_block(n=5, cell=0.25, ground=True),
        _warm(scenes.box_stack_3d(levels=3, drop=0.02), 10),
        scenes.cloth_sheet(n=6, spacing=0.25),
    ]),
}


@pytest.mark.parametrize("lanes", [(1, 1), (8, 16)], ids=["seq", "wide"])
@pytest.mark.parametrize("case", list(CASES))
def test_graph_replay_matches_eager(case, lanes):
    world = CASES[case]()
    steps = 12
    with Context(lanes=lanes[0], fem_lanes=lanes[1]) as g, Context(lanes=lanes[0], fem_lanes=lanes[1]) as e:
        e.set_option("step_graphs", 0)
        g.upload(world)
        e.upload(world)
        for k in range(steps):
            g.step(1)
            e.step(1)
            _equal(g.download(), e.download(), (case, k))
        replays, _ = g.kernel_stats("step_graph")
        eager_replays, _ = e.kernel_stats("step_graph")
    assert eager_replays == 0
    # the settled pyramid keeps its shape from step to step
    if case == "pyramid_settled":
        assert replays >= steps // 2, replays


def test_graph_replay_with_scene_subsets():
    """Steps over changing subsets of scenes (the served form's ticks) replay
    graphs whose kernels skip the idle scenes at run time."""
    worlds = [_warm(scenes.pyramid_2d(rows=5), 20), scenes.joint_chain(5), scenes.fem_block(n=4, cell=0.25)]
    world = World.concat(worlds)
    subsets = [[0, 1, 2], [0, 2], [0, 1, 2], [1], [0, 1, 2], [0, 2], [0, 1, 2], [0, 2]]
    with Context(lanes=4, fem_lanes=4) as g, Context(lanes=4, fem_lanes=4) as e:
        e.set_option("step_graphs", 0)
        g.upload(world)
        e.upload(world)
        for k, ids in enumerate(subsets):
            g.step_scenes(ids, 1)
            e.step_scenes(ids, 1)
            _equal(g.download(), e.download(), k)
        replays, captures = g.kernel_stats("step_graph")
    assert replays >= 1 and captures >= 1, (replays, captures)
