"""Short horizons as diagnostics: the GPU and the FP32 oracle (colour-major,
the order the GPU computes) advanced k steps from the same state, each
compared with the FP64 oracle.  Single steps are held to the parity bars
(test_gpu_parity.py); over many steps contact-rich scenes are chaotic, so
here the GPU's distance from FP64 is compared with the FP32 oracle's own
distance from FP64: both carry FP32 rounding, and the GPU must not drift
further than a small multiple of what FP32 arithmetic itself does.  Smooth
scenes (joints, springs, FEM, cloth) also keep the GPU within a fixed bound
of the FP32 oracle.  Every horizon is printed (and appended to $PARITY_OUT).
"""
import json
import os

import numpy as np
import pytest

import oracle_py as O
from parity import pos_rel_err, rot_err
from skyephyx_amd import scenes
from skyephyx_amd.capi import Context

pytestmark = pytest.mark.gpu

HORIZONS = (1, 10, 30, 60)

# smooth: no contacts (or resting ones that do not switch), errors stay small
SMOOTH = {
    "joint_chain": lambda: scenes.joint_chain(5),
    "spring_bridge": lambda: scenes.spring_bridge(6),
    "fem_block": lambda: scenes.fem_block(n=6, cell=0.25),
}
# contact-rich or unstable: piles that settle, and the cloth sheet, whose
# particles fold onto each other (the FP32 and FP64 oracles themselves end
# ~0.35 m apart per metre after 60 steps); chaotic over long horizons
CHAOTIC = {
    "cloth": lambda: scenes.cloth_sheet(n=8, spacing=0.25),
    "pyramid": lambda: scenes.pyramid_2d(),
    "box_stack_3d": lambda: scenes.box_stack_3d(levels=3, drop=0.02),
    "lattice_mix": lambda: scenes.box_lattice(shape=(4, 3, 3), spacing=3.0, n_chain=12, n_springs=8, bounds=50.0),
}


def _horizons(make):
    world = make()
    gpu = Context(lanes=1, fem_lanes=1)  # the reference's accumulation order
    gpu.upload(world)
    o32 = O.Oracle(world, fp64=False, order=O.COLOUR_MAJOR)
    o64 = O.Oracle(world, fp64=True, order=O.COLOUR_MAJOR)
    done = 0
    rows = []
    for k in HORIZONS:
        gpu.step(k - done)
        o32.step(k - done)
        o64.step(k - done)
        done = k
        g, a, d = gpu.download(), o32.save(), o64.save()
        if len(g.bodies) != len(a.bodies) or len(a.bodies) != len(d.bodies):
            rows.append({"steps": k, "bodies_differ": True})
            break
        rows.append({
            "steps": k,
            "gpu_vs_fp32": pos_rel_err(g.bodies, a.bodies),
            "gpu_vs_fp64": pos_rel_err(g.bodies, d.bodies),
            "fp32_vs_fp64": pos_rel_err(a.bodies, d.bodies),
            "rot_gpu_vs_fp64": rot_err(g.bodies, d.bodies),
            "rot_fp32_vs_fp64": rot_err(a.bodies, d.bodies),
        })
    gpu.close()
    return rows


def _report(name, rows):
    print(name)
    for r in rows:
        print("  ", json.dumps(r))
    out = os.environ.get("PARITY_OUT")
    if out:
        with open(out, "a") as f:
            f.write(json.dumps({"horizon_case": name, "rows": rows}) + "\n")


@pytest.mark.parametrize("case", sorted(SMOOTH))
def test_short_horizon_smooth(case):
    rows = _horizons(SMOOTH[case])
    _report(case, rows)
    for r in rows:
        assert not r.get("bodies_differ"), r
        # within a fixed bound of the FP32 oracle, and no further from FP64
        # than a few times the FP32 oracle is
        assert r["gpu_vs_fp32"] <= 1e-3, r
        assert r["gpu_vs_fp64"] <= 4.0 * r["fp32_vs_fp64"] + 1e-5, r


@pytest.mark.parametrize("case", sorted(CHAOTIC))
def test_short_horizon_contact_rich(case):
    rows = _horizons(CHAOTIC[case])
    _report(case, rows)
    assert rows[0]["gpu_vs_fp32"] <= 1e-4, rows[0]  # one step: the parity bar
    for r in rows:
        assert not r.get("bodies_differ"), r
        assert np.isfinite(r["gpu_vs_fp64"]), r
        assert r["gpu_vs_fp64"] <= 10.0 * r["fp32_vs_fp64"] + 1e-3, r
