"""Vectorized scene generators for large worlds (FEM blocks with 10^5-10^7
tets, box lattices with 10^4-10^6 boxes, the 4096-scene batched mix).

They produce the same records as WorldBuilder (the reference's construction
API restated in world.py) would, built with numpy in bulk; a test checks the
two agree on small sizes.
"""
from __future__ import annotations

import math

import numpy as np

from .world import (
    BODY_DTYPE,
    BODY_PARTICLE,
    BODY_STATIC,
    F32,
    JOINT_DTYPE,
    SPRING_DTYPE,
    TET_DTYPE,
    World,
    WorldBuilder,
    default_scene,
    MANIFOLD_DTYPE,
    TRI_DTYPE,
    SCENE_DTYPE,
)


def _radius(scale):
    s = scale.astype(F32)
    d = (s[:, 0] * s[:, 0] + s[:, 1] * s[:, 1]) + s[:, 2] * s[:, 2]
    return np.sqrt(d).astype(F32) * F32(0.5)


def boxes(pos, scale, rot=None, density=1.0, friction=0.5, static=None):
    """Solver::AddBody for cubes, vectorized."""
    n = len(pos)
    b = np.zeros(n, BODY_DTYPE)
    sc = np.asarray(scale, F32).reshape(n, 3)
    b["pos"] = np.asarray(pos, F32)
    b["rot"] = (1.0, 0.0, 0.0, 0.0) if rot is None else np.asarray(rot, F32)
    b["scale"] = sc
    b["friction"] = friction
    b["radius"] = _radius(sc)
    st = np.zeros(n, bool) if static is None else np.asarray(static, bool)
    dens = F32(density)
    sx, sy, sz = sc[:, 0], sc[:, 1], sc[:, 2]
    m = ((dens * sx) * sy) * sz
    k = m / F32(12.0)
    inertia = np.stack([k * (sy * sy + sz * sz), k * (sx * sx + sz * sz), k * (sx * sx + sy * sy)], axis=1).astype(F32)
    b["mass"] = np.where(st, F32(0), m)
    b["inertia"] = np.where(st[:, None], F32(0), inertia)
    b["flags"] = np.where(st, BODY_STATIC, 0).astype(np.uint32)
    return b


def particles(pos, mass, friction=0.5, static=None):
    """Solver::AddParticle, vectorized (0.1 m cube, zero inertia)."""
    n = len(pos)
    s = F32(0.1)
    density = F32(mass) / ((s * s) * s)
    b = boxes(pos, np.full((n, 3), s, F32), density=density, friction=friction, static=static)
    b["flags"] = b["flags"] | BODY_PARTICLE
    b["inertia"] = 0.0
    return b


def tets(positions, idx, E, nu, serial0=0):
    """NeoHookeanFEM ctor, vectorized: positions (V,3) float32, idx (T,4)."""
    T = len(idx)
    t = np.zeros(T, TET_DTYPE)
    t["v"] = idx
    t["serial"] = serial0 + np.arange(T)
    p = positions.astype(np.float64)
    p0 = p[idx[:, 0]]
    Dm = np.stack([p[idx[:, 1]] - p0, p[idx[:, 2]] - p0, p[idx[:, 3]] - p0], axis=2)
    t["dm_inv"] = np.linalg.inv(Dm).astype(F32).reshape(T, 9)
    t["rest_volume"] = (np.abs(np.linalg.det(Dm)) / 6.0).astype(F32)
    mu, lam = WorldBuilder.lame(E, nu)
    t["mu"] = mu
    t["lambda"] = lam
    return t


def grid_tets(n):
    """Connectivity of makeTetVolume (geometry.cpp:59-93) for an (nx,ny,nz)
    vertex grid with vertex index i + j*nx + k*nx*ny, in the reference's
    creation order (ci outer, then cj, then ck; 5 tets per cell)."""
    nx, ny, nz = n
    ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()

    def vid(i, j, k):
        return i + j * nx + k * nx * ny

    v = [vid(ci, cj, ck), vid(ci + 1, cj, ck), vid(ci, cj + 1, ck), vid(ci + 1, cj + 1, ck),
         vid(ci, cj, ck + 1), vid(ci + 1, cj, ck + 1), vid(ci, cj + 1, ck + 1), vid(ci + 1, cj + 1, ck + 1)]
    pattern = [(0, 1, 2, 4), (3, 2, 1, 7), (5, 1, 4, 7), (6, 4, 2, 7), (1, 2, 4, 7)]
    out = np.stack([np.stack([v[a], v[b], v[c], v[d]], axis=1) for a, b, c, d in pattern], axis=1)
    return out.reshape(-1, 4).astype(np.int32)


def fem_block(n=32, cell=0.25, E=1e5, nu=0.3, rho=1000.0, pin_top=True, origin=None, **scene):
    """Same world as scenes.fem_block, built in bulk."""
    nx = ny = nz = n
    if origin is None:
        origin = (-(n - 1) * cell * 0.5, 0.0, -(n - 1) * cell * 0.5)
    ox, oy, oz = origin
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    pos = np.stack([ox + i * cell, oy + j * cell, oz + k * cell], axis=1).astype(F32)
    static = (j == n - 1) if pin_top else np.zeros(len(i), bool)
    b = particles(pos, rho * cell ** 3, static=static)
    t = tets(b["pos"], grid_tets((nx, ny, nz)), E, nu)
    return World(np.array([default_scene(**scene)], SCENE_DTYPE), b, np.zeros(0, JOINT_DTYPE),
                 np.zeros(0, SPRING_DTYPE), np.zeros(0, MANIFOLD_DTYPE), t, np.zeros(0, TRI_DTYPE))


def random_quaternions(rng, n):
    """Uniform unit quaternions (Shoemake 1992), (w, x, y, z)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, c = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.stack([c * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  c * np.sin(2 * np.pi * u3)], axis=1)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(F32)


def box_pile(n_boxes, rng, spacing=1.9, edge=(0.4, 1.0), **scene):
    """n_boxes randomly sized and oriented boxes on a lattice above a static
    ground box (as levels.h:30, widened with the footprint).  The footprint
    grows with the box count so piles stay about 2 layers tall, and the
    lattice pitch exceeds the largest bounding-sphere diameter (sqrt(3) x the
    largest edge) so no boxes overlap initially."""
    cols = max(2, int(math.ceil(math.sqrt(n_boxes / 2.0))))
    per_layer = cols * cols
    idx = np.arange(n_boxes)
    layer = idx // per_layer
    r = idx % per_layer
    ix, iz = r % cols, r // cols
    # the first layer's bounding spheres (radius <= 0.87) clear the ground top (y = -9.5)
    pos = np.stack([(ix - (cols - 1) * 0.5) * spacing, -9.5 + 0.9 + layer * spacing, (iz - (cols - 1) * 0.5) * spacing],
                   axis=1)
    pos = pos + rng.uniform(-0.02, 0.02, pos.shape)
    scale = rng.uniform(edge[0], edge[1], (n_boxes, 3))
    rot = random_quaternions(rng, n_boxes)
    side = max(20.0, cols * spacing + 16.0)
    ground = boxes(np.array([[0.0, -10.0, 0.0]]), np.array([[side, 1.0, side]]), static=[True])
    b = np.concatenate([ground, boxes(pos, scale, rot)])
    return World(np.array([default_scene(**scene)], SCENE_DTYPE), b, np.zeros(0, JOINT_DTYPE),
                 np.zeros(0, SPRING_DTYPE), np.zeros(0, MANIFOLD_DTYPE), np.zeros(0, TET_DTYPE),
                 np.zeros(0, TRI_DTYPE))


def batched_mix(n_scenes=4096, seed=7, box_range=(8, 2048), tet_range=(1000, 50000), fem_fraction=0.5):
    """The batched workload: independent scenes of uneven size.  Rigid scenes
    hold a log-uniform number of boxes in box_range; FEM scenes are hanging
    Neo-Hookean blocks (top face pinned) with a log-uniform tet count in
    tet_range; each scene has dt in {1/60, 1/120} and 8-16 iterations."""
    rng = np.random.default_rng(seed)
    worlds = []
    for s in range(n_scenes):
        dt = F32(1.0) / F32(60.0 if rng.random() < 0.5 else 120.0)
        iters = int(rng.integers(8, 17))
        if rng.random() >= fem_fraction:
            nb = int(round(math.exp(rng.uniform(math.log(box_range[0]), math.log(box_range[1])))))
            worlds.append(box_pile(nb, rng, dt=dt, iterations=iters))
        else:
            nt = math.exp(rng.uniform(math.log(tet_range[0]), math.log(tet_range[1])))
            cells = max(1, int(round((nt / 5.0) ** (1.0 / 3.0))))
            worlds.append(fem_block(n=cells + 1, cell=0.25, dt=dt, iterations=iters))
    return World.concat(worlds)
