"""Parity at the full sizes of the benchmark configurations.

Each world is advanced on the GPU into the state the configuration is about
(a sagging block, a folding sheet, a landed pile, a settling mix), downloaded,
and then one step is taken from that identical state by the GPU and by the
oracle (colour-major order, what the GPU computes; FP64 beside it; the
reference's natural order reported), under the single-step bars of
tests/parity.py:

* FEM block: a 32^3-vertex grid split 5 tets per cube (148,955 Neo-Hookean
  tets, E = 1e5, nu = 0.3, rho = 1000, top face pinned), and the StVK
  variant, a 32 x 32 membrane sheet (the reference's StVK is a triangle
  membrane element, energy.hpp:88-121);
* 20k boxes with edges U[0.5, 2] m in a jittered 40 x 25 x 20 lattice, a
  1k-link ball-joint chain and 5k random springs (k = 1e4), after the boxes
  have landed;
* the batched mix: 4096 scenes (box piles of 8-2048 boxes, Neo-Hookean
  blocks of 1k-50k tets, dt in {1/60, 1/120}, 8-16 iterations) stepped as
  one world, a size-stratified sample of scenes each compared with the
  oracle stepping that scene alone;
* the served form over the same mix, every scene resident under a 200 GB
  budget: sampled scenes are bit-identical to the scene stepped alone.
"""
import json
import os
import time

import numpy as np
import pytest

from parity import assert_parity, one_step_report
from skyephyx_amd import scenes, scenes_fast
from skyephyx_amd.capi import Context, footprint
from skyephyx_amd.serve import Server

pytestmark = pytest.mark.gpu


def _record(name, res):
    out = os.environ.get("PARITY_OUT")
    print(name, json.dumps(res, default=str))
    if out:
        with open(out, "a") as f:
            f.write(json.dumps({"case": name, **res}, default=str) + "\n")


class _Progress:
    """Live progress lines (pytest captures test output; a stage of these
    tests can run for tens of seconds)."""

    def __init__(self, capsys, name):
        self.capsys, self.name, self.t0 = capsys, name, time.perf_counter()

    def __call__(self, msg):
        with self.capsys.disabled():
            print(f"[{self.name} {time.perf_counter() - self.t0:6.1f}s] {msg}", flush=True)


def _gpu_warm(world, steps, say, lanes=(4, 4)):
    """Advance on the GPU with fixed lane widths (the autotuner's choice
    depends on timings, so a tuned warm-up would not give the same state on
    every run)."""
    with Context(lanes=lanes[0], fem_lanes=lanes[1]) as ctx:
        ctx.upload(world)
        for k in range(0, steps, 50):
            ctx.step(min(50, steps - k))
        ctx.synchronize()
        say(f"warmed {steps} steps on the GPU")
        return ctx.download()


def _gpu_one_step(state, lanes):
    with Context(lanes=lanes[0], fem_lanes=lanes[1]) as ctx:
        ctx.upload(state)
        ctx.step(1)
        g = ctx.download()
        return g, ctx.colours(len(g.bodies))


def _single_step_case(name, state, lanes, say):
    g, gcol = _gpu_one_step(state, lanes)
    say(f"GPU step lanes {lanes}; oracle steps ...")
    res = one_step_report(state, g, gcol)
    res.update(n_bodies=int(len(state.bodies)), n_manifolds_before=int(len(state.manifolds)),
               n_tets=int(len(state.tets)), n_tris=int(len(state.tris)))
    _record(f"{name}/lanes{lanes[0]}-{lanes[1]}", res)
    say(f"pos {res['pos_rel']:.2e} rot {res['rot']:.2e} lambda(norm) {res['lambda_norm_rel']:.2e}")
    assert_parity(res, strict_lambda=lanes == (1, 1))


@pytest.mark.parametrize("lanes", [(1, 1), (0, 0)], ids=["seq", "tuned"])
def test_fem_block_32(lanes, capsys):
    say = _Progress(capsys, "fem32")
    world = scenes_fast.fem_block(32, cell=0.25)
    assert len(world.tets) == 148955 and len(world.bodies) == 32 ** 3
    state = _gpu_warm(world, 20, say)
    _single_step_case("fem_block_32", state, lanes, say)


@pytest.mark.parametrize("lanes", [(1, 1), (0, 0)], ids=["seq", "tuned"])
def test_stvk_sheet_32(lanes, capsys):
    say = _Progress(capsys, "sheet32")
    world = scenes.cloth_sheet(n=32, spacing=0.25)
    assert len(world.tris) == 2 * 31 * 31
    state = _gpu_warm(world, 20, say)
    _single_step_case("stvk_sheet_32", state, lanes, say)


@pytest.mark.parametrize("lanes", [(1, 1), (0, 0)], ids=["seq", "tuned"])
def test_boxes_20k_landed(lanes, capsys):
    say = _Progress(capsys, "boxes20k")
    world = scenes.box_lattice(shape=(40, 25, 20), spacing=3.5, n_chain=1000, n_springs=5000)
    assert len(world.bodies) == 1 + 20000 + 1000 and len(world.joints) == 1000 and len(world.springs) == 5000
    # the top layer starts ~84 m above the ground: 6 s of fall and settling
    state = _gpu_warm(world, 360, say)
    dyn = (state.bodies["flags"] & 1) == 0
    vy = np.abs(state.bodies["vel"][dyn, 1])
    v = np.abs(state.bodies["vel"]).max(axis=1)
    say(f"{len(state.manifolds)} manifolds, median |vy| {np.median(vy):.3f} m/s, max |v| boxes "
        f"{v[1:20001].max():.1f} chain {v[20001:].max():.1f} m/s")
    assert state.scenes["emergency_stop"][0] == 0 and len(state.bodies) == len(world.bodies)
    assert len(state.manifolds) > 10000  # in contact, not in free fall
    _single_step_case("boxes_20k_landed", state, lanes, say)


@pytest.fixture(scope="module")
def mix():
    return scenes_fast.batched_mix(4096)


def _stratified(world, per_kind=8):
    """Scene ids: per kind (box piles, FEM blocks) `per_kind` scenes evenly
    spaced over the size order."""
    n_tets = np.bincount(world.bodies["scene"][world.tets["v"][:, 0]], minlength=len(world.scenes))
    n_bodies = np.bincount(world.bodies["scene"], minlength=len(world.scenes))
    out = []
    for fem in (False, True):
        ids = np.nonzero((n_tets > 0) == fem)[0]
        order = ids[np.argsort((n_tets if fem else n_bodies)[ids], kind="stable")]
        pick = np.unique(np.round(np.linspace(0, len(order) - 1, per_kind)).astype(int))
        out.extend(int(order[k]) for k in pick)
    return out


def test_batched_mix_4096_scenes_against_the_oracle_per_scene(mix, capsys):
    """The 4096-scene world stepped as one (autotuned lanes, step graphs, the
    product's defaults); sampled scenes (both kinds, smallest to largest,
    both time steps, several iteration counts) against the oracle stepping
    that scene alone from the same state."""
    say = _Progress(capsys, "mix4096")
    sample = _stratified(mix)
    assert len(sample) >= 16
    dts = set(np.round(1.0 / mix.scenes["dt"][sample]).astype(int).tolist())
    iters = set(mix.scenes["iterations"][sample].tolist())
    assert dts == {60, 120} and len(iters) >= 4, (dts, iters)
    with Context() as ctx:
        ctx.upload(mix)
        ctx.synchronize()
        say(f"uploaded {len(mix.scenes)} scenes, {len(mix.bodies)} bodies, {len(mix.tets)} tets")
        ctx.step(30)  # piles in contact
        ctx.synchronize()
        pre = ctx.download()
        say(f"warmed 30 steps, {len(pre.manifolds)} manifolds; downloaded")
        ctx.step(1)
        post = ctx.download()
        gcol = ctx.colours(len(post.bodies))
        say("stepped once more and downloaded")
    for s in sample:
        state = pre.extract_scene(s)
        g = post.extract_scene(s)
        res = one_step_report(state, g, gcol[post.scene_bodies(s)], natural=False)
        res.update(scene=s, n_bodies=int(len(state.bodies)), n_tets=int(len(state.tets)),
                   n_manifolds_before=int(len(state.manifolds)), dt=float(state.scenes["dt"][0]),
                   iterations=int(state.scenes["iterations"][0]))
        _record(f"mix4096/scene{s}", res)
        assert_parity(res)
    say(f"{len(sample)} sampled scenes within the bars")


def test_served_mix_4096_scenes_alone_equals_mixed(mix, capsys):
    """Every scene of the mix resident under a 200 GB budget, requests of
    1-4 steps for every scene arriving in two waves (continuous batching:
    the second wave joins while the first runs); sampled scenes after their
    requests are bit-identical to the scene stepped alone for the same
    number of steps (the server's accumulation rules: 4 lanes per body with
    forces, canonical FEM sums; the scene alone runs its FEM particles at 4
    lanes, the server at whatever width its tuner picks per launch)."""
    say = _Progress(capsys, "served4096")
    assert footprint(mix) <= 200e9
    rng = np.random.default_rng(11)
    steps = rng.integers(1, 5, len(mix.scenes))
    sample = _stratified(mix)
    with Server(budget_bytes=200e9) as srv:
        ids = list(srv.add_scenes(mix))
        srv.preload()
        st = srv.stats()
        assert st["resident_scenes"] == len(mix.scenes), st
        say(f"preloaded {len(ids)} scenes, {st['resident_bytes'] / 1e9:.1f} GB resident")
        half = len(ids) // 2
        for i in ids[:half]:
            srv.submit(i, int(steps[i]))
        srv.tick()
        for i in ids[half:]:
            srv.submit(i, int(steps[i]))
        srv.run_until_idle()
        done = srv.poll()
        assert len(done) == len(ids)
        served = {s: srv.download_scene(ids[s]) for s in sample}
        say("served every request; sampled scenes downloaded")
    for s in sample:
        one = mix.extract_scene(s)
        with Context(lanes=4) as ctx:
            ctx.set_option("fem_canonical_sums", 1)
            ctx.set_option("fem_lanes", 4)
            ctx.upload(one)
            ctx.step(int(steps[s]))
            ref = ctx.download()
        g = served[s]
        for f in ("pos", "rot", "vel", "ang_vel"):
            assert np.array_equal(g.bodies[f], ref.bodies[f]), (s, f)
        assert len(g.manifolds) == len(ref.manifolds), s
        for f in ("body_a", "body_b", "n_contacts"):
            assert np.array_equal(g.manifolds[f], ref.manifolds[f]), (s, f)
        for f in ("id", "lambda", "penalty"):
            assert np.array_equal(g.manifolds["contacts"][f], ref.manifolds["contacts"][f]), (s, f)
    say(f"{len(sample)} sampled scenes bit-identical alone and served")
