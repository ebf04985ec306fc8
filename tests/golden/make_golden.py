"""Writes the golden fixtures: small worlds and their state after a few steps
of the FP32 oracle (natural visit order).  They pin regressions of the
oracle and of the scene builders; they are not reference outputs (the
reference ships none and cannot be built here).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle_py as O  # noqa: E402
from skyephyx_amd import scenes  # noqa: E402
from skyephyx_amd.world import World  # noqa: E402

FIELDS = ("scenes", "bodies", "joints", "springs", "manifolds", "tets", "tris")


def cases():
    return {
        "pyramid_c1": (scenes.pyramid_2d(), [1, 10, 40]),
        "tet_block_2x2x2": (scenes.fem_block(n=2, cell=0.5), [1, 5]),
        "cloth_2x2": (scenes.cloth_sheet(n=2, spacing=0.5), [1, 5]),
        "joint_pair": (scenes.joint_chain(2), [1, 10]),
        "spring_pair": (scenes.mass_springs(2), [1, 10]),
        "spring_bridge": (scenes.spring_bridge(3), [1, 30]),
    }


def save(path, w: World):
    np.savez_compressed(path, **{k: getattr(w, k) for k in FIELDS})


def _current(a, dtype):
    """Records of an earlier layout in the current dtype, field by field
    (fields added since are zero: the reference's behaviour)."""
    if a.dtype == dtype:
        return a
    out = np.zeros(a.shape, dtype)
    for name in a.dtype.names:
        out[name] = a[name]
    return out


def load(path) -> World:
    from skyephyx_amd.abi import _DTYPES

    with np.load(path, allow_pickle=False) as z:
        return World(*(_current(z[k], dt) for k, dt in zip(FIELDS, _DTYPES)))


def main():
    for name, (w, steps) in cases().items():
        save(os.path.join(HERE, f"{name}_in.npz"), w)
        o = O.Oracle(w, fp64=False, order=O.NATURAL)
        done = 0
        for s in steps:
            o.step(s - done)
            done = s
            save(os.path.join(HERE, f"{name}_step{s}.npz"), o.save())
        print(name, "ok")


if __name__ == "__main__":
    main()
