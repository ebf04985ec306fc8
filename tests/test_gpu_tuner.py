"""The launch-shape autotuner (lanes per body, then workgroup size, per
kernel class and launch-size band) on the GPU: it converges on live
launches, its table round-trips through avbd_tune_table / avbd_tune_load
(also the earlier width-only format), a context seeded with a table samples
nothing (its step graphs start at once), the workgroup size never changes a
result, and a step with loaded shapes is within the parity bars of the
oracle."""
import numpy as np
import pytest

import oracle_py as O
from parity import assert_parity, one_step_report
from skyephyx_amd import scenes, scenes_fast
from skyephyx_amd.capi import Context
from skyephyx_amd.world import World

pytestmark = pytest.mark.gpu


def _world():
    o = O.Oracle(scenes_fast.box_pile(300, np.random.default_rng(2)), order=O.NATURAL)
    o.step(20)
    return World.concat([o.save(), scenes.fem_block(n=10, cell=0.25), scenes.cloth_sheet(n=8, spacing=0.25)])


def _tuned_table(world, steps=60):
    with Context() as ctx:
        ctx.upload(world)
        ctx.step(steps)
        return ctx.tune_table(), ctx.download()


def test_tuner_converges_and_round_trips():
    world = _world()
    table, _ = _tuned_table(world)
    assert table, "nothing tuned"
    for (cls, band), (lanes, block) in table.items():
        assert cls in (0, 1) and 0 <= band < 31
        assert lanes in (1, 2, 4, 8, 16, 32) and block in (64, 128, 256)
    replays = []
    for seed in (False, True):
        with Context() as ctx:
            if seed:
                ctx.tune_load(table)
                assert ctx.tune_table() == table
            ctx.upload(world)
            ctx.kernel_stats("step_graph", reset=True)
            ctx.step(12)
            replays.append(ctx.kernel_stats("step_graph")[0])
            if seed:  # what was loaded is kept (keys it lacked may have been added)
                t = ctx.tune_table()
                assert all(t[k] == v for k, v in table.items())
    # a step graph needs a launch shape without a key being sampled: the seeded
    # context reaches that sooner than a fresh one
    assert replays[1] > replays[0], replays
    # the earlier format: a bare width takes workgroup size 256
    with Context() as ctx:
        ctx.tune_load({k: v[0] for k, v in table.items()})
        assert ctx.tune_table() == {k: (v[0], 256) for k, v in table.items()}


def test_workgroup_size_never_changes_results():
    world = _world()
    table, _ = _tuned_table(world, steps=40)
    outs = []
    for block in (64, 128, 256):
        with Context() as ctx:
            ctx.tune_load({k: (v[0], block) for k, v in table.items()})
            ctx.upload(world)
            ctx.step(6)
            outs.append(ctx.download())
    for g in outs[1:]:
        assert g.bodies.tobytes() == outs[0].bodies.tobytes()
        assert g.manifolds.tobytes() == outs[0].manifolds.tobytes()


def test_loaded_shapes_within_parity_bars():
    world = _world()
    table, state = _tuned_table(world, steps=40)
    with Context() as ctx:
        ctx.tune_load(table)
        ctx.upload(state)
        ctx.step(1)
        g = ctx.download()
        gcol = ctx.colours(len(g.bodies))
    for s in range(len(state.scenes)):
        idx = g.scene_bodies(s)
        res = one_step_report(state.extract_scene(s), g.extract_scene(s), gcol[idx], natural=False)
        assert_parity(res)
