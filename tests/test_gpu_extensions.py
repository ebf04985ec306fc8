"""GPU parity of the two features the reference declares but leaves
unfinished (scene options; off, the reference's behaviour, by default):
sphere collision and the trust-region ratio of ADAPTIVE projection.

One step from an identical state, GPU (C ABI) against the oracle in
colour-major order, under the single-step bars of tests/parity.py (contact
pairs, counts, feature ids and colours bit-exact; positions 1e-4; lambda
1e-3); the trust-region ratio the step leaves within 1e-3 of the oracle's.
With an option off the GPU step is bit-identical to the step of the
reference's behaviour (spheres as boxes, ADAPTIVE as ABSOLUTE).
"""
import json

import numpy as np
import pytest

import oracle_py as O
from parity import assert_parity, one_step_report
from skyephyx_amd import scenes, scenes_fast
from skyephyx_amd.capi import Context
from skyephyx_amd.world import PROJ_ABSOLUTE, PROJ_ADAPTIVE, PROJ_CLAMP, SCENE_TRUST_REGION

pytestmark = pytest.mark.gpu


def _gpu_run(world, steps, lanes=(4, 4), options=()):
    with Context(lanes=lanes[0], fem_lanes=lanes[1]) as ctx:
        for k, v in options:
            ctx.set_option(k, v)
        ctx.upload(world)
        if steps:
            ctx.step(steps)
        return ctx.download()


def _gpu_one_step(state, lanes, options=()):
    with Context(lanes=lanes[0], fem_lanes=lanes[1]) as ctx:
        for k, v in options:
            ctx.set_option(k, v)
        ctx.upload(state)
        ctx.step(1)
        g = ctx.download()
        return g, ctx.colours(len(g.bodies))


@pytest.mark.parametrize("lanes", [(1, 1), (4, 4)], ids=["seq", "lanes4"])
@pytest.mark.parametrize("warm", [0, 45, 90])
def test_sphere_pile_matches_oracle(lanes, warm):
    """Spheres and boxes falling on the ground and onto each other
    (sphere-sphere, sphere-box, sphere-ground contacts)."""
    state = _gpu_run(scenes.sphere_pile(), warm)
    g, gcol = _gpu_one_step(state, lanes)
    res = one_step_report(state, g, gcol, natural=False)
    res.update(n_manifolds_before=int(len(state.manifolds)), warm=warm)
    print("sphere_pile", lanes, json.dumps(res, default=str))
    assert_parity(res, strict_lambda=lanes == (1, 1))
    if warm:
        kinds = {int(c["id"]) & 3 for m in state.manifolds for c in m["contacts"][: m["n_contacts"]]}
        assert 3 in kinds, kinds  # sphere contacts were live


def test_sliding_sphere_matches_oracle_along_its_run():
    state = scenes.sliding_sphere(v0=4.0, radius=0.5, friction=0.3)
    for k in range(4):
        g, gcol = _gpu_one_step(state, (1, 1))
        res = one_step_report(state, g, gcol, natural=False)
        assert_parity(res, strict_lambda=True)
        state = _gpu_run(g, 10, lanes=(1, 1))


def test_spheres_collide_as_boxes_without_the_option():
    w = scenes.sphere_pile(options=0)
    boxes = w.copy()
    boxes.bodies["flags"] = boxes.bodies["flags"] & ~np.uint32(4)
    a = _gpu_run(w, 60)
    b = _gpu_run(boxes, 60)
    assert a.bodies["pos"].tobytes() == b.bodies["pos"].tobytes()
    assert a.bodies["rot"].tobytes() == b.bodies["rot"].tobytes()


def _fem(E, rho0, mode=PROJ_ADAPTIVE, options=SCENE_TRUST_REGION, n=8):
    return scenes_fast.fem_block(n=n, cell=0.25, E=E, projection_mode=mode, options=options, trust_region_rho=rho0)


@pytest.mark.parametrize("variant", ["lanes1", "lanes4", "canonical"])
def test_trust_region_step_matches_oracle(variant):
    lanes = (1, 1) if variant == "lanes1" else (4, 4)
    opts = (("fem_canonical_sums", 1),) if variant == "canonical" else ()
    state = _gpu_run(_fem(1e5, 0.0), 20, lanes=lanes, options=opts)
    g, gcol = _gpu_one_step(state, lanes, opts)
    res = one_step_report(state, g, gcol, natural=False)
    rho = {}
    for fp64 in (False, True):
        o = O.Oracle(state, fp64=fp64, order=O.COLOUR_MAJOR)
        o.step(1)
        rho[fp64] = float(o.save().scenes["trust_region_rho"][0])
    rho_gpu = float(g.scenes["trust_region_rho"][0])
    res.update(rho_gpu=rho_gpu, rho_oracle=rho[False], rho_oracle64=rho[True],
               rho_before=float(state.scenes["trust_region_rho"][0]))
    print("trust_region", variant, json.dumps(res, default=str))
    assert_parity(res)
    # the ratio within the FP32 oracle's own distance from FP64 (x3), at least 1e-3
    bar = max(1e-3, 3.0 * abs(rho[False] - rho[True]))
    assert rho_gpu != 0.0 and abs(rho_gpu - rho[False]) <= bar, (rho_gpu, rho, bar)


def test_trust_region_rho_reported_by_the_abi():
    with Context(lanes=4, fem_lanes=4) as ctx:
        ctx.upload(_fem(1e5, 0.0))
        ctx.step(3)
        rho = ctx.trust_region_rho(1)
        assert float(rho[0]) == float(ctx.download().scenes["trust_region_rho"][0]) and 0.5 < rho[0] < 1.5


def test_adaptive_without_the_option_is_absolute():
    a = _gpu_run(_fem(1e5, 0.0, options=0), 10)
    b = _gpu_run(_fem(1e5, 0.0, mode=PROJ_ABSOLUTE, options=0), 10)
    assert float(a.scenes["trust_region_rho"][0]) == 0.0
    assert a.bodies["pos"].tobytes() == b.bodies["pos"].tobytes()


def test_adaptive_near_one_equals_clamp():
    """A stiff block started at rho = 1 keeps |rho - 1| <= 0.01 (checked), so
    each sweep projects by CLAMP: bit-identical to a CLAMP run."""
    with Context(lanes=4, fem_lanes=4) as a, Context(lanes=4, fem_lanes=4) as c:
        a.upload(_fem(1e7, 1.0))
        c.upload(_fem(1e7, 0.0, mode=PROJ_CLAMP, options=0))
        for _ in range(3):
            a.step(1)
            c.step(1)
            assert abs(float(a.trust_region_rho(1)[0]) - 1.0) < 1e-3
            assert a.download().bodies["pos"].tobytes() == c.download().bodies["pos"].tobytes()
