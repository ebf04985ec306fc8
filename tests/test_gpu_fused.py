"""Per-scene kernel on the GPU: small scenes run phases 4-7 (prediction,
iterations, velocities, post-stabilisation) in one workgroup each instead of
the per-colour bucket launches.  Same per-body and per-force functions in
the same order, so every step must be bit-identical to the bucket path
(option fused_scenes 1 against 0, with the single-scene default off), also
for bounds removal (post-stabilisation after the removal runs in the same
kernel) and in worlds that mix small scenes, large scenes and scenes with
breakable forces (which stay on the buckets).  Scenes whose bodies and
incidence lists fit run on an LDS copy of them; the 300-box pile does not
fit and runs on global memory.
"""
import numpy as np
import pytest

import oracle_py as O
from skyephyx_amd import scenes, scenes_fast
from skyephyx_amd.capi import Context
from skyephyx_amd.world import World

pytestmark = pytest.mark.gpu


def _warm(world, steps):
    o = O.Oracle(world, fp64=False, order=O.NATURAL)
    o.step(steps)
    return o.save()


def _equal(a: World, b: World, where):
    assert len(a.bodies) == len(b.bodies), where
    for f in ("pos", "rot", "vel", "ang_vel", "prev_vel"):
        assert np.array_equal(a.bodies[f], b.bodies[f]), (where, f)
    assert len(a.manifolds) == len(b.manifolds), where
    for f in ("body_a", "body_b", "serial", "n_contacts"):
        assert np.array_equal(a.manifolds[f], b.manifolds[f]), (where, f)
    for f in ("lambda", "penalty", "stick", "id"):
        assert np.array_equal(a.manifolds["contacts"][f], b.manifolds["contacts"][f]), (where, f)
    assert np.array_equal(a.joints["lambda"], b.joints["lambda"]), where
    assert np.array_equal(a.springs["lambda"], b.springs["lambda"]), where
    assert np.array_equal(a.scenes["emergency_stop"], b.scenes["emergency_stop"]), where
    assert np.array_equal(a.scenes["n_removed"], b.scenes["n_removed"]), where


CASES = {
    "pyramid": lambda: scenes.pyramid_2d(),
    "pyramid_settled": lambda: _warm(scenes.pyramid_2d(), 40),
    "box_stack": lambda: _warm(scenes.box_stack_3d(levels=3, drop=0.02), 10),
    "joint_chain": lambda: scenes.joint_chain(6),
    "mass_springs": lambda: scenes.mass_springs(4),
    "cloth": lambda: scenes.cloth_sheet(n=6, spacing=0.25),
    "fem_block": lambda: scenes.fem_block(n=5, cell=0.25),
    "out_of_bounds": lambda: scenes.out_of_bounds(),
    "pile_over_lds": lambda: _warm(scenes_fast.box_pile(300, np.random.default_rng(2)), 30),
    "mixed_world": lambda: World.concat([
        _warm(scenes.pyramid_2d(rows=6), 20),
        scenes.spring_bridge(6),          # breakable springs: bucket path
        scenes.fem_block(n=9, cell=0.25),  # 729 vertices: above the size limit, bucket path
        scenes.joint_chain(5),
        scenes.out_of_bounds(),
        scenes.cloth_sheet(n=5, spacing=0.25),
    ]),
}


@pytest.mark.parametrize("lanes", [1, 4], ids=["seq", "g4"])
@pytest.mark.parametrize("case", list(CASES))
def test_per_scene_kernel_matches_buckets(case, lanes):
    world = CASES[case]()
    with Context(lanes=lanes, fem_lanes=lanes) as f, Context(lanes=lanes, fem_lanes=lanes) as b:
        f.set_option("fused_scenes", 1)
        b.set_option("fused_scenes", 0)
        b.set_option("fused_single_scene", 0)
        f.upload(world)
        b.upload(world)
        for k in range(8):
            f.step(1)
            b.step(1)
            _equal(f.download(), b.download(), (case, k))


def test_per_scene_kernel_with_scene_subsets():
    """Served-style ticks over changing scene subsets."""
    world = World.concat([_warm(scenes.pyramid_2d(rows=5), 20), scenes.joint_chain(5), scenes.fem_block(n=4, cell=0.25),
                          scenes.spring_bridge(4)])
    subsets = [[0, 1, 2, 3], [0, 2], [1, 3], [0, 1, 2, 3], [2], [0, 1]]
    with Context(lanes=4, fem_lanes=4) as f, Context(lanes=4, fem_lanes=4) as b:
        f.set_option("fused_scenes", 1)
        b.set_option("fused_scenes", 0)
        b.set_option("fused_single_scene", 0)
        f.upload(world)
        b.upload(world)
        for k, ids in enumerate(subsets):
            f.step_scenes(ids, 1)
            b.step_scenes(ids, 1)
            _equal(f.download(), b.download(), k)


def test_scene_lds_copy_matches_global_memory():
    """The per-scene kernel on LDS copies (default) and on global memory
    (option scene_lds_bytes 0) produce the same bits; the option is range
    checked."""
    world = CASES["mixed_world"]()
    with Context(lanes=4, fem_lanes=4) as l, Context(lanes=4, fem_lanes=4) as g:
        l.set_option("fused_scenes", 1)
        g.set_option("fused_scenes", 1)
        g.set_option("scene_lds_bytes", 0)
        with pytest.raises(Exception, match="scene_lds_bytes"):
            g.set_option("scene_lds_bytes", 65537)
        l.upload(world)
        g.upload(world)
        for k in range(6):
            l.step(1)
            g.step(1)
            _equal(l.download(), g.download(), k)
