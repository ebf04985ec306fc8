"""FEM-only sweeps with neighbour lists (option fem_neighbour_lists, on by
default): a vertex's lane group loads its distinct element neighbours'
positions into LDS once and every element reads them from there.  The
positions are the values the per-element gathers read, so the results are
bit-identical to the sweep without lists: checked over several steps on
Neo-Hookean blocks, a StVK sheet, a block in contact with the ground, a
removal inside FEM scenes, and a vertex with more neighbours than a list
holds (it falls back to the gathers)."""
import numpy as np
import pytest

from skyephyx_amd import scenes
from skyephyx_amd.capi import Context
from skyephyx_amd.world import MODEL_CUBE, World, WorldBuilder

pytestmark = pytest.mark.gpu


def _fan(n_tets=24):
    """A vertex shared by n_tets tets whose other vertices are all distinct:
    3 n_tets neighbours, more than a list holds."""
    b = WorldBuilder(gravity=(0.0, -9.81, 0.0))
    c = b.add_particle(0.1, 0.5, (0.0, 0.0, 0.0))
    for i in range(n_tets):
        a = 2 * np.pi * i / n_tets
        d = (np.cos(a), 0.0, np.sin(a))
        # radial, radial + up, radial + tangential: a proper tet with the centre
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
                                     scenes.fem_block(n=8, cell=0.25)]),
    "fan": _fan,
}


@pytest.mark.parametrize("lanes", [4, 8, 16], ids=["g4", "g8", "g16"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_neighbour_lists_bit_identical(case, lanes):
    world = CASES[case]()
    outs = []
    for on in (1, 0):
        with Context(lanes=lanes, fem_lanes=lanes) as ctx:
            ctx.set_option("fem_neighbour_lists", on)
            ctx.upload(world)
            ctx.step(12)
            outs.append(ctx.download())
    a, b = outs
    assert a.bodies.tobytes() == b.bodies.tobytes()
    assert a.manifolds.tobytes() == b.manifolds.tobytes()


def test_neighbour_lists_after_removal():
    """Removing bodies renumbers a scene's vertices and drops the elements
    that lost one; the scene's lists are rebuilt from its new streams."""
    world = World.concat([scenes.fem_block(n=6, cell=0.25), scenes.fem_block(n=5, cell=0.25)])
    outs = []
    for on in (1, 0):
        with Context(lanes=4, fem_lanes=4) as ctx:
            ctx.set_option("fem_neighbour_lists", on)
            ctx.upload(world)
            ctx.step(3)
            ctx.remove_bodies([7, 40, 250])
            ctx.step(6)
            outs.append(ctx.download())
    assert outs[0].bodies.tobytes() == outs[1].bodies.tobytes()
    assert len(outs[0].tets) == len(outs[1].tets) < len(world.tets)
