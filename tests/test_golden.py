"""The oracle and the scene builders reproduce the committed golden fixtures."""
import glob
import os
import re

import numpy as np
import pytest

import oracle_py as O
from golden.make_golden import cases, load

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture_cases():
    out = []
    for path in sorted(glob.glob(os.path.join(HERE, "*_in.npz"))):
        name = os.path.basename(path)[: -len("_in.npz")]
        steps = sorted(int(re.search(r"_step(\d+)\.npz", p).group(1))
                       for p in glob.glob(os.path.join(HERE, f"{name}_step*.npz")))
        out.append((name, steps))
    return out


def test_fixtures_exist():
    assert len(fixture_cases()) == len(cases())


@pytest.mark.parametrize("name", sorted(cases()))
def test_builders_reproduce_inputs(name):
    w, _ = cases()[name]
    ref = load(os.path.join(HERE, f"{name}_in.npz"))
    assert w.bodies.tobytes() == ref.bodies.tobytes()
    assert w.joints.tobytes() == ref.joints.tobytes()
    assert w.springs.tobytes() == ref.springs.tobytes()
    assert w.tets.tobytes() == ref.tets.tobytes()
    assert w.tris.tobytes() == ref.tris.tobytes()


@pytest.mark.parametrize("name,steps", fixture_cases())
def test_oracle_reproduces_fixture(name, steps):
    w = load(os.path.join(HERE, f"{name}_in.npz"))
    o = O.Oracle(w, fp64=False, order=O.NATURAL)
    done = 0
    for s in steps:
        o.step(s - done)
        done = s
        got = o.save()
        ref = load(os.path.join(HERE, f"{name}_step{s}.npz"))
        np.testing.assert_allclose(got.bodies["pos"], ref.bodies["pos"], rtol=0, atol=1e-5)
        np.testing.assert_allclose(got.bodies["rot"], ref.bodies["rot"], rtol=0, atol=1e-5)
        assert len(got.manifolds) == len(ref.manifolds)
        if len(ref.manifolds):
            np.testing.assert_array_equal(got.manifolds["contacts"]["id"], ref.manifolds["contacts"]["id"])
