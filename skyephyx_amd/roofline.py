"""Algorithmic-byte model of the AVBD sweeps (compulsory structure-of-arrays
traffic per visit, no cross-visit reuse), used for roofline fractions.

Per primal block update of one body:
  rigid body      140 B (x, q, y, q_y, m, world inertia, x_last, q_last read;
                         x, q written)
                  + 60 B per incident manifold (other body's x, q, x_last,
                         q_last, friction)
                  + 108 B per incident contact point (one side's 3x6 Jacobian,
                         C0, lambda, penalty)
                  + 228 B per incident joint, + 88 B per incident spring
  FEM vertex       40 B (x read/write, y, m) + 104 B per incident tet/triangle
                         (3 other x, rest-shape inverse, V0/mu/lambda, ids,
                         list entry)
Per dual pass: 212 B per contact point, 116 B per manifold, 76 B per joint or
spring row.
"""
from __future__ import annotations

import numpy as np

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, GB/s

RIGID_BASE, PER_MANIFOLD, PER_CONTACT, PER_JOINT, PER_SPRING = 140, 60, 108, 228, 88
FEM_BASE, PER_ELEMENT = 40, 104
DUAL_CONTACT, DUAL_MANIFOLD, DUAL_ROW = 212, 116, 76


def sweep_bytes(world):
    """Per-body bytes of one primal visit and the number of visits per step.

    Returns (bytes_per_visit[N], visits_per_step[N], dual_bytes_per_iter[S],
    iterations[S])."""
    b = world.bodies
    N = len(b)
    dyn = (b["flags"] & 1) == 0
    n_man = np.zeros(N, np.int64)
    n_con = np.zeros(N, np.int64)
    for side in ("body_a", "body_b"):
        if len(world.manifolds):
            np.add.at(n_man, world.manifolds[side], 1)
            np.add.at(n_con, world.manifolds[side], world.manifolds["n_contacts"])
    n_joint = np.zeros(N, np.int64)
    n_spring = np.zeros(N, np.int64)
    for arr, cnt in ((world.joints, n_joint), (world.springs, n_spring)):
        if len(arr):
            a = arr["body_a"]
            np.add.at(cnt, a[a >= 0], 1)
            np.add.at(cnt, arr["body_b"], 1)
    n_elem = np.zeros(N, np.int64)
    if len(world.tets):
        np.add.at(n_elem, world.tets["v"].ravel(), 1)
    if len(world.tris):
        np.add.at(n_elem, world.tris["v"].ravel(), 1)
    base = np.where(n_elem > 0, FEM_BASE, RIGID_BASE)
    per = (base + PER_MANIFOLD * n_man + PER_CONTACT * n_con + PER_JOINT * n_joint + PER_SPRING * n_spring +
           PER_ELEMENT * n_elem)
    sc = world.scenes
    visits_scene = sc["iterations"].astype(np.int64) + (sc["post_stabilization"] != 0)
    visits = np.where(dyn, visits_scene[b["scene"]], 0)
    S = len(sc)
    dual = np.zeros(S, np.int64)
    if len(world.manifolds):
        s_of = b["scene"][world.manifolds["body_a"]]
        np.add.at(dual, s_of, DUAL_MANIFOLD + DUAL_CONTACT * world.manifolds["n_contacts"].astype(np.int64))
    if len(world.joints):
        np.add.at(dual, b["scene"][world.joints["body_b"]], 6 * DUAL_ROW)
    if len(world.springs):
        np.add.at(dual, b["scene"][world.springs["body_b"]], DUAL_ROW)
    return per.astype(np.int64), visits.astype(np.int64), dual, sc["iterations"].astype(np.int64)


def step_bytes(world):
    """Algorithmic bytes of one step: (primal, dual)."""
    per, visits, dual, iters = sweep_bytes(world)
    return int((per * visits).sum()), int((dual * iters).sum())


def fem_only_mask(world):
    """Bodies the primal sweep puts in its FEM-only bucket: particles with
    element energies and no contact point, joint or spring."""
    b = world.bodies
    N = len(b)
    forces = np.zeros(N, np.int64)
    if len(world.manifolds):
        for side in ("body_a", "body_b"):
            np.add.at(forces, world.manifolds[side], world.manifolds["n_contacts"])
    for arr in (world.joints, world.springs):
        if len(arr):
            a = arr["body_a"]
            np.add.at(forces, a[a >= 0], 1)
            np.add.at(forces, arr["body_b"], 1)
    n_elem = np.zeros(N, np.int64)
    if len(world.tets):
        np.add.at(n_elem, world.tets["v"].ravel(), 1)
    if len(world.tris):
        np.add.at(n_elem, world.tris["v"].ravel(), 1)
    particle = np.all(b["inertia"] == 0, axis=1)
    return (forces == 0) & (n_elem > 0) & particle


def class_step_bytes(world):
    """Primal algorithmic bytes of one step split by bucket:
    (bodies with forces, FEM-only particles)."""
    per, visits, _, _ = sweep_bytes(world)
    fem = fem_only_mask(world)
    tot = per * visits
    return int(tot[~fem].sum()), int(tot[fem].sum())


def body_updates_per_step(world):
    """Primal block updates per step (dynamic bodies x (iterations + post-stab))."""
    _, visits, _, _ = sweep_bytes(world)
    return int(visits.sum())
