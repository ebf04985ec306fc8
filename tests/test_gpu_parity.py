"""GPU-vs-oracle parity of one AVBD step from an identical state.

The HIP path runs through the C ABI; the oracle runs the same step in
colour-major visit order (bodies of one colour share no force, so this is
exactly what the per-colour parallel sweep computes).  Bars: positions and
orientations within 1e-4 relative, lambda within 1e-3 relative; contact
pairs, counts, feature IDs and colours bit-exact.  The FP64 oracle error of
the same step is reported beside it.
"""
import json
import os

import numpy as np
import pytest

import oracle_py as O
from skyephyx_amd import scenes, scenes_fast
from skyephyx_amd.capi import Context
from parity import LAMBDA_TOL, POS_TOL, summary

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
    accumulates in the reference's sequential order."""
    make, warm = CASES[case]
    state = warm_state(make(), warm)
    ctx = Context(lanes=lanes[0], fem_lanes=lanes[1])
    ctx.upload(state)
    ctx.step(1)
    g = ctx.download()
    gcol = ctx.colours(len(g.bodies))
    o = O.Oracle(state, fp64=False, order=O.COLOUR_MAJOR)
    o.step(1)
    ref = o.save()
    ocol, _ = o.colours(len(ref.bodies))
    d = O.Oracle(state, fp64=True, order=O.COLOUR_MAJOR)
    d.step(1)
    ref64 = d.save()
    res = summary(g, ref)
    dyn = (g.bodies["flags"] & 1) == 0
    res["colours_equal"] = bool(np.array_equal(gcol[dyn], ocol[dyn]))
    res["fp64_pos_rel_gpu"] = summary(g, ref64)["pos_rel"]
    spread = summary(ref, ref64)  # the FP32 oracle against the FP64 oracle
    res["fp64_pos_rel_oracle32"] = spread["pos_rel"]
    res["oracle32_vs_64_lambda_rel"] = spread["lambda_rel"]
    _record(f"{case}/lanes{lanes[0]}-{lanes[1]}", res)
    assert res["pairs_equal"], res
    assert res["n_equal"] and res["ids_equal"], res
    assert res["colours_equal"], res
    assert res["pos_rel"] <= POS_TOL, res
    assert res["rot"] <= POS_TOL, res
    # lambda: 1e-3 relative over all contact rows; per entry, no further from
    # the FP32 oracle than twice the FP32 oracle's own distance from FP64
    assert res["lambda_norm_rel"] <= LAMBDA_TOL, res
    assert res["lambda_rel"] <= max(LAMBDA_TOL, 2.0 * spread["lambda_rel"]), res
    if "joint_lambda_rel" in res:
        assert res["joint_lambda_rel"] <= LAMBDA_TOL, res
    if "spring_lambda_rel" in res:
        assert res["spring_lambda_rel"] <= LAMBDA_TOL, res
