"""CPU-side checks: the C-ABI library loads and exports what the header
declares, record layouts agree, world construction and generators agree,
and the byte model is consistent.  No GPU compute here."""
import subprocess
import sys

import numpy as np
import pytest

from skyephyx_amd import capi, roofline, scenes, scenes_fast, world
from skyephyx_amd.abi import AvbdRequestInfo, AvbdServerStats, AvbdTimings, AvbdWorld
import ctypes as C


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    assert L.avbd_abi_version() == 2
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = capi.exported_symbols()
    assert len(declared) >= 15
    missing = [s for s in declared if s not in exported]
    assert not missing, missing


def test_record_layouts_match_numpy_dtypes():
    sizes = (C.c_int32 * 16)()
    n = capi.lib().avbd_record_sizes(sizes, 16)
    assert n == 12
    assert list(sizes)[:8] == [d.itemsize for d in world.RECORD_DTYPES]
    assert sizes[8] == C.sizeof(AvbdWorld)
    assert sizes[9] == C.sizeof(AvbdTimings)
    assert sizes[10] == C.sizeof(AvbdRequestInfo)
    assert sizes[11] == C.sizeof(AvbdServerStats)


def test_footprint_is_host_only_and_consistent():
    """avbd_footprint needs no device.  The device world of several scenes
    shares its fixed-size arrays, so it is smaller than the sum of the
    scenes' worlds, and it grows with every element kind."""
    a = scenes.box_stack_3d(levels=3)
    b = scenes.fem_block(n=4, cell=0.25)
    c = scenes.joint_chain(5)
    fa, fb, fc = capi.footprint(a), capi.footprint(b), capi.footprint(c)
    both = capi.footprint(world.World.concat([a, b, c]))
    assert 0 < max(fa, fb, fc) < both < fa + fb + fc
    bigger = capi.footprint(scenes.box_stack_3d(levels=4))
    assert bigger > fa
    assert capi.footprint(a, manifold_capacity=1 << 16) > capi.footprint(a, manifold_capacity=1 << 12)


def test_server_without_gpu_fails_loudly():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except ImportError:
        pass
    from skyephyx_amd.serve import Server

    with pytest.raises(capi.AvbdError):
        Server()


def test_create_without_gpu_fails_loudly():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except ImportError:
        pass
    with pytest.raises(capi.AvbdError):
        capi.Context()


def test_builder_constructor_semantics():
    b = world.WorldBuilder()
    i = b.add_body(world.MODEL_CUBE, 2.0, 0.5, (0, 0, 0), (1.0, 2.0, 3.0))
    r = b.body(i)
    assert r["mass"] == np.float32(12.0)
    np.testing.assert_allclose(r["inertia"], [12 / 12 * 13, 12 / 12 * 10, 12 / 12 * 5], rtol=1e-6)
    np.testing.assert_allclose(r["radius"], np.sqrt(14.0) / 2, rtol=1e-6)
    p = b.add_particle(0.5, 0.5, (1, 1, 1))
    assert b.body(p)["flags"] & world.BODY_PARTICLE and np.all(b.body(p)["inertia"] == 0)
    np.testing.assert_allclose(b.body(p)["mass"], 0.5, rtol=1e-6)
    s = b.add_body(world.MODEL_CUBE, 1.0, 0.5, (0, 5, 0), (1, 1, 1), is_static=True)
    assert b.body(s)["mass"] == 0 and b.body(s)["flags"] & world.BODY_STATIC
    j = b.add_joint(None, (0, 6, 0), i, (0, 0.5, 0))
    np.testing.assert_allclose(b._joints[j]["rest_rot"], [1, 0, 0, 0])
    np.testing.assert_allclose(b._joints[j]["torque_arm"], 1 + 4 + 9)
    sp = b.add_spring(s, (0, 0, 0), i, (0, 0, 0), 100.0)
    np.testing.assert_allclose(b.spring(sp)["rest_length"], 5.0)
    with pytest.raises(ValueError):
        b.add_body(world.MODEL_PYRAMID, 1.0, 0.5, (0, 0, 0), (1, 1, 1))


def test_fast_generators_match_builder():
    a = scenes.fem_block(n=4, cell=0.25)
    b = scenes_fast.fem_block(n=4, cell=0.25)
    assert a.bodies.tobytes() == b.bodies.tobytes()
    assert a.tets.tobytes() == b.tets.tobytes()


def test_concat_and_extract_round_trip():
    w1 = scenes.pyramid_2d(rows=3)
    w2 = scenes.joint_chain(3)
    w = world.World.concat([w1, w2])
    assert len(w.scenes) == 2
    back = w.extract_scene(1)
    assert back.bodies.tobytes() == w2.bodies.tobytes()
    assert np.array_equal(back.joints["body_b"], w2.joints["body_b"])


def test_byte_model_counts():
    w = scenes.fem_block(n=3, cell=0.25, iterations=10)
    per, visits, dual, iters = roofline.sweep_bytes(w)
    nt = np.zeros(len(w.bodies), int)
    np.add.at(nt, w.tets["v"].ravel(), 1)
    dyn = (w.bodies["flags"] & 1) == 0
    np.testing.assert_array_equal(per, 40 + 104 * nt)
    assert np.all(visits[dyn] == 11) and np.all(visits[~dyn] == 0)
    assert roofline.body_updates_per_step(w) == 11 * int(dyn.sum())


def test_bench_rejects_more_than_one_gpu():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
