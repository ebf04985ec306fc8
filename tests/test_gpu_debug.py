"""The renderer's line and point data made on the device (option
debug_output) against the oracle's restatement of the reference's
AddLineData / AddDebugPointData calls during the step (solver.cpp:266-313;
manifold.cpp:259-302, joint.cpp:176-190, spring.cpp:112-125,
neohookeanfem.cpp:113-133, stvkfem.cpp:89-105).  The device groups the lines
by kind where the reference follows its force and energy vectors, so the
two are compared as sets: every line (point) is matched one to one with a
line (point) of the same colour, within 1e-5 m after one step from the same
state and within the position parity bar (1e-4 relative to max(|x|, 1 m))
after a second step, whose start poses already carry the first step's FP32
differences.
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


CASES = {
    "pyramid_settled": lambda: _warm(scenes.pyramid_2d(), 40),  # contacts: lines and points
    "box_stack_3d": lambda: _warm(scenes.box_stack_3d(levels=3, drop=0.02), 10),
    "joint_chain": lambda: scenes.joint_chain(5),  # incl. a world anchor
    "spring_bridge": lambda: scenes.spring_bridge(6),
    "fem_block": lambda: scenes.fem_block(n=4, cell=0.25),
    "cloth": lambda: scenes.cloth_sheet(n=5, spacing=0.25),
    "lattice_mix": lambda: _warm(scenes.box_lattice(shape=(3, 3, 2), spacing=3.0, n_chain=6, n_springs=5,
                                                    bounds=50.0), 25),
}


def _match(a, b, what, rel=0.0):
    """One-to-one match of the rows of a and b: same colour, geometry within
    1e-5, or rel * max(|x|, 1) (greedy over the nearest candidates)."""
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if len(a) == 0:
        return
    ng = a.shape[1] - 4  # geometry columns before the colour
    used = np.zeros(len(b), bool)
    for row in a:
        d = np.abs(b[:, :ng] - row[:ng]).max(axis=1)
        d[used] = np.inf
        d[np.any(b[:, ng:] != row[ng:], axis=1)] = np.inf
        k = int(np.argmin(d))
        tol = max(1e-5, rel * max(1.0, float(np.abs(row[:ng]).max())))
        assert d[k] <= tol, (what, row, b[k], d[k])
        used[k] = True


@pytest.mark.parametrize("case", sorted(CASES))
def test_debug_output_matches_the_oracle(case):
    world = CASES[case]()
    o = O.Oracle(world, fp64=False, order=O.COLOUR_MAJOR)
    with Context(lanes=1, fem_lanes=1) as ctx:
        ctx.set_option("debug_output", 1)
        ctx.upload(world)
        for k in range(2):
            ctx.step(1)
            o.step(1)
            gl, gp = ctx.debug_output()
            ol, op = o.debug()
            rel = 0.0 if k == 0 else 1e-4
            _match(gl, ol, (case, k, "lines"), rel)
            _match(gp, op, (case, k, "points"), rel)


def test_debug_output_colours_and_scenes():
    """Element edges take the mean colour of their bodies; several scenes
    output together; off by default."""
    world = World.concat([scenes.fem_block(n=3, cell=0.5), scenes.joint_chain(3)])
    n = len(world.bodies)
    rgb = np.zeros((n, 3), np.float32)
    rgb[:, 0] = np.linspace(0.0, 1.0, n)
    rgb[:, 2] = 0.5
    with Context(lanes=1, fem_lanes=1) as ctx:
        ctx.upload(world)
        ctx.step(1)
        lines, points = ctx.debug_output()
        assert len(lines) == 0 and len(points) == 0  # option off
        ctx.set_option("debug_output", 1)
        ctx.set_body_colours(rgb)
        ctx.step(1)
        lines, points = ctx.debug_output()
    tets = world.tets["v"]
    assert len(lines) == 6 * len(tets) + len(world.joints)
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    want = np.array([(rgb[t[i]] + rgb[t[j]]) * 0.5 for t in tets for i, j in edges], np.float32)
    np.testing.assert_allclose(lines[len(world.joints):, 6:9], want, rtol=0, atol=1e-7)
    assert np.all(lines[:, 9] == 1.0)


def _rotmat(q):
    w, x, y, z = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_instances_match_model_and_normal_matrices():
    """avbd_get_instances: per live body (dead slots skipped, after a removal)
    the model matrix [R diag(scale) | t] column-major and the normal matrix
    as the reference computes it, inverse(3x3 block) transposed (numpy's
    general inverse here), with the body's colour."""
    world = _warm(scenes.box_lattice(shape=(3, 3, 2), spacing=3.0, n_chain=4, n_springs=3, bounds=50.0), 10)
    with Context(lanes=4, fem_lanes=4) as ctx:
        ctx.upload(world)
        ctx.remove_bodies([4, 9])
        n = len(world.bodies) - 2
        rgb = np.random.default_rng(1).random((n, 3)).astype(np.float32)
        ctx.set_body_colours(rgb)
        ctx.step(2)
        inst = ctx.instances()
        bodies = ctx.download().bodies
    assert inst.shape == (n, 32)
    for i, b in enumerate(bodies):
        RS = _rotmat(b["rot"]) @ np.diag(b["scale"].astype(np.float64))
        model = np.eye(4)
        model[:3, :3] = RS
        model[:3, 3] = b["pos"]
        np.testing.assert_allclose(inst[i, :16].reshape(4, 4).T, model, rtol=0, atol=2e-6 * max(1.0, np.abs(b["pos"]).max()))
        normal = np.linalg.inv(RS).T
        got = inst[i, 16:28].reshape(3, 4)  # three columns of 4
        np.testing.assert_allclose(got[:, :3].T, normal, rtol=0, atol=1e-5 * np.abs(normal).max())
        assert np.all(got[:, 3] == 0)
        np.testing.assert_array_equal(inst[i, 28:31], rgb[i])
        assert inst[i, 31] == 1.0
