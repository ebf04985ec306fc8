"""avbd_get_instances_noalloc (no device allocation; its staging lives in
the library's per-GPU scratch array) against avbd_get_instances, bit for
bit, over calls in a row whose inputs shrink, across two contexts on one
GPU, and its refusal of worlds beyond the scratch's size."""
import numpy as np
import pytest

from skyephyx_amd import scenes, scenes_fast
from skyephyx_amd.capi import NOALLOC_MAX_BODIES, AvbdError, Context

pytestmark = pytest.mark.gpu


def _same(ctx):
    ref = ctx.instances()
    got = ctx.instances(noalloc=True)
    assert got.shape == ref.shape
    assert got.tobytes() == ref.tobytes()
    return len(ref)


def test_noalloc_instances_equal_the_allocating_entry_point():
    big = scenes_fast.box_pile(3000, np.random.default_rng(5))
    small = scenes.box_lattice(shape=(3, 3, 2), spacing=3.0, n_chain=4, n_springs=3, bounds=50.0)
    with Context(lanes=4, fem_lanes=4) as a, Context(lanes=4, fem_lanes=4) as b:
        a.upload(big)
        a.remove_bodies([7, 100, 2500])  # dead slots are skipped
        a.set_body_colours(np.random.default_rng(2).random((len(big.bodies) - 3, 3)).astype(np.float32))
        a.step(3)
        n1 = _same(a)
        assert n1 == len(big.bodies) - 3
        # second call with a smaller input: the scratch still holds the
        # first call's rows beyond it
        a.upload(small)
        a.step(2)
        n2 = _same(a)
        assert n2 == len(small.bodies) < n1
        # another context on the same GPU shares the scratch (calls in turn)
        b.upload(scenes.fem_block(n=4, cell=0.5))
        b.step(1)
        _same(b)
        _same(a)
        # a short output buffer gets the first cap instances; *n is the count
        import ctypes as C
        out = np.zeros((5, 32), np.float32)
        n = C.c_int32()
        assert a.L.avbd_get_instances_noalloc(a.h, out.ctypes.data, 5, C.byref(n)) == 0
        assert n.value == n2
        assert out.tobytes() == a.instances()[:5].tobytes()


def test_noalloc_instances_refuse_worlds_beyond_the_scratch():
    from skyephyx_amd.world import World

    two = World.concat([scenes.joint_chain(3), scenes.joint_chain(3)])
    with Context() as ctx:
        # body slots = bodies + 2 x 65536 spare slots > the scratch's 131072
        ctx.set_option("spare_bodies", NOALLOC_MAX_BODIES // 2)
        ctx.upload(two)
        with pytest.raises(AvbdError, match="at most 131072"):
            ctx.instances(noalloc=True)
        assert len(ctx.instances()) == len(two.bodies)
