"""FEM-only sweeps with canonical lane sums (option fem_canonical_sums): a
particle's entries are shared out over 32 virtual lanes and summed by one
fixed tree whatever the launch width, so 4, 8, 16 and 32 lanes, and the
tuner's own choices, give bit-identical worlds.  Checked over several steps
on Neo-Hookean blocks, a StVK sheet, a block in contact with the ground (the
paired forces + FEM launch), a vertex with more than 32 entries, and across
a removal; one step from warm states against the FP32 oracle as in the
parity suite; option validation."""
import numpy as np
import pytest

from skyephyx_amd import scenes
from skyephyx_amd.capi import AvbdError, Context
from skyephyx_amd.world import World, WorldBuilder

pytestmark = pytest.mark.gpu

from parity import assert_parity, one_step_report  # noqa: E402
from test_gpu_parity import warm_state  # noqa: E402


def _fan(n_tets=40):
    """A vertex shared by n_tets tets (more entries than the 32 virtual lanes)."""
    b = WorldBuilder(gravity=(0.0, -9.81, 0.0))
    c = b.add_particle(0.1, 0.5, (0.0, 0.0, 0.0))
    for i in range(n_tets):
        a = 2 * np.pi * i / n_tets
        d = (np.cos(a), 0.0, np.sin(a))
        q = [(0.3 * d[0], 0.0, 0.3 * d[2]), (0.3 * d[0], 0.3, 0.3 * d[2]),
             (0.3 * d[0] - 0.2 * d[2], 0.0, 0.3 * d[2] + 0.2 * d[0])]
        p = [b.add_particle(0.1, 0.5, q[k], is_static=k == 1) for k in range(3)]
        b.add_tet(c, p[0], p[1], p[2], 1e4, 0.3)
    return b.build()


CASES = {
    "fem_block": lambda: scenes.fem_block(n=7, cell=0.25),
    "cloth": lambda: scenes.cloth_sheet(n=10, spacing=0.25),
    "fem_on_ground": lambda: scenes.fem_block(n=6, cell=0.25, pin_top=False, ground=True, origin=(-0.6, -9.45, -0.6)),
    "several": lambda: World.concat([scenes.fem_block(n=5, cell=0.25), scenes.cloth_sheet(n=6, spacing=0.25),
                                     scenes.fem_block(n=8, cell=0.25), scenes.pyramid_2d()]),
    "fan": _fan,
}


def _run(world, fem_lanes, steps=12, removal=None):
    with Context(lanes=4) as ctx:
        ctx.set_option("fem_canonical_sums", 1)
        ctx.set_option("fem_lanes", fem_lanes)
        ctx.upload(world)
        if removal:
            ctx.step(3)
            ctx.remove_bodies(removal)
        ctx.step(steps)
        return ctx.download()


@pytest.mark.parametrize("case", sorted(CASES))
def test_canonical_sums_width_independent(case):
    world = CASES[case]()
    ref = _run(world, 4)
    for g in (8, 16, 32, 0):
        out = _run(world, g)
        assert out.bodies.tobytes() == ref.bodies.tobytes(), (case, g)
        assert out.manifolds.tobytes() == ref.manifolds.tobytes(), (case, g)


def test_canonical_sums_after_removal():
    world = World.concat([scenes.fem_block(n=6, cell=0.25), scenes.fem_block(n=5, cell=0.25)])
    ref = _run(world, 4, steps=6, removal=[7, 40, 250])
    for g in (8, 32):
        out = _run(world, g, steps=6, removal=[7, 40, 250])
        assert out.bodies.tobytes() == ref.bodies.tobytes(), g
    assert len(ref.tets) < len(world.tets)


PARITY = {
    "fem_block": (lambda: scenes.fem_block(n=7, cell=0.25), 5),
    "cloth": (lambda: scenes.cloth_sheet(n=10, spacing=0.25), 5),
    "fem_on_ground": (lambda: scenes.fem_block(n=9, cell=0.25, pin_top=False, ground=True,
                                               origin=(-1.0, -9.45, -1.0)), 5),
}


@pytest.mark.parametrize("fem_lanes", [4, 32])
@pytest.mark.parametrize("case", sorted(PARITY))
def test_canonical_sums_parity(case, fem_lanes):
    make, warm = PARITY[case]
    state = warm_state(make(), warm)
    with Context(lanes=4) as ctx:
        ctx.set_option("fem_canonical_sums", 1)
        ctx.set_option("fem_lanes", fem_lanes)
        ctx.upload(state)
        ctx.step(1)
        g = ctx.download()
        gcol = ctx.colours(len(g.bodies))
    assert_parity(one_step_report(state, g, gcol, natural=False))


def test_canonical_sums_option_checks():
    with Context() as ctx:
        ctx.set_option("fem_lanes", 2)
        with pytest.raises(AvbdError):
            ctx.set_option("fem_canonical_sums", 1)
        ctx.set_option("fem_lanes", 0)
        ctx.set_option("fem_canonical_sums", 1)
        with pytest.raises(AvbdError):
            ctx.set_option("fem_lanes", 1)
        ctx.set_option("fem_lanes", 16)
