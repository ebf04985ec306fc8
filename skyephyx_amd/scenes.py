"""Seeded scene generators for the benchmark and test workloads.

They build worlds through WorldBuilder, i.e. with the reference's
construction API.  The FEM block follows makeTetVolume's connectivity
(src/helpers/geometry.cpp:30-94, 5 tets per cell), the membrane follows
ClothSimulation (src/levels.h:442-491), the joint chain follows the pendulum
of JointPlayground (levels.h:203-233) and the ground box is the one of every
level (levels.h:30).
"""
from __future__ import annotations

import math

import numpy as np

from .world import (
    F32,
    INF,
    MODEL_CUBE,
    MODEL_SPHERE,
    SCENE_SPHERE_COLLISION,
    World,
    WorldBuilder,
)

GROUND_POS = (0.0, -10.0, 0.0)
GROUND_SCALE = (20.0, 1.0, 20.0)


def add_ground(b: WorldBuilder, scale=GROUND_SCALE, pos=GROUND_POS):
    return b.add_body(MODEL_CUBE, 1.0, 0.5, pos, scale, is_static=True)


def pyramid_2d(rows=10, gap=0.1, lift=0.02, **scene):
    """Rows of unit boxes (rows, rows-1, ..., 1) centred at x=0, each row offset
    half a pitch from the one below and `lift` metres above it; the bottom row
    is `lift` above the ground top (y = -9.5)."""
    params = dict(alpha=0.99, gamma=0.99, beta=1e5)
    params.update(scene)
    b = WorldBuilder(**params)
    add_ground(b)
    pitch = 1.0 + gap
    for r in range(rows):
        n = rows - r
        y = -9.5 + lift + 0.5 + r * (1.0 + lift)
        for k in range(n):
            x = (k - (n - 1) * 0.5) * pitch
            b.add_body(MODEL_CUBE, 1.0, 0.5, (x, y, 0.0), (1.0, 1.0, 1.0))
    return b.build()


def box_stack_3d(levels=4, spacing=0.3, drop=0.0, **scene):
    """The reference's 3-D Pyramid level (levels.h:28-82), parameterised."""
    b = WorldBuilder(**scene)
    add_ground(b)
    scale = (2.0, 1.0, 2.0)
    for row in range(levels):
        nps = levels - row
        span = nps * scale[0] + (nps - 1) * spacing
        sx = -span * 0.5 + scale[0] * 0.5
        sz = -span * 0.5 + scale[2] * 0.5
        y = -9.5 + scale[1] * 0.5 + row * (scale[1] + spacing) + drop
        for ix in range(nps):
            for iz in range(nps):
                b.add_body(MODEL_CUBE, 1.0, 0.5, (sx + ix * (scale[0] + spacing), y, sz + iz * (scale[2] + spacing)),
                           scale)
    return b.build()


def tet_volume(b: WorldBuilder, origin, n, cell, E, nu, particle_mass, friction=0.5, pin=None):
    """makeTetVolume (geometry.cpp:30-94).  `pin(i, j, k)` marks static vertices."""
    nx, ny, nz = n
    ox, oy, oz = origin
    idx = np.zeros((nx, ny, nz), np.int64)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                st = bool(pin(i, j, k)) if pin else False
                idx[i, j, k] = b.add_particle(particle_mass, friction, (ox + i * cell, oy + j * cell, oz + k * cell),
                                              is_static=st)
    for ci in range(nx - 1):
        for cj in range(ny - 1):
            for ck in range(nz - 1):
                v = [idx[ci, cj, ck], idx[ci + 1, cj, ck], idx[ci, cj + 1, ck], idx[ci + 1, cj + 1, ck],
                     idx[ci, cj, ck + 1], idx[ci + 1, cj, ck + 1], idx[ci, cj + 1, ck + 1], idx[ci + 1, cj + 1, ck + 1]]
                b.add_tet(v[0], v[1], v[2], v[4], E, nu)
                b.add_tet(v[3], v[2], v[1], v[7], E, nu)
                b.add_tet(v[5], v[1], v[4], v[7], E, nu)
                b.add_tet(v[6], v[4], v[2], v[7], E, nu)
                b.add_tet(v[1], v[2], v[4], v[7], E, nu)
    return idx


def fem_block(n=32, cell=0.25, E=1e5, nu=0.3, rho=1000.0, pin_top=True, origin=None, ground=False, **scene):
    """n^3-vertex Neo-Hookean block, top face (j = n-1) pinned, hanging in gravity."""
    b = WorldBuilder(**scene)
    if ground:
        add_ground(b)
    if origin is None:
        origin = (-(n - 1) * cell * 0.5, 0.0, -(n - 1) * cell * 0.5)
    mass = rho * cell ** 3
    tet_volume(b, origin, (n, n, n), cell, E, nu, mass, pin=(lambda i, j, k: j == n - 1) if pin_top else None)
    return b.build()


def cloth_sheet(n=32, m=None, spacing=0.25, E=1e4, nu=0.3, mass=0.1, y=5.0, **scene):
    """STVK membrane sheet as in ClothSimulation (levels.h:442-491): two corners pinned."""
    m = n if m is None else m
    b = WorldBuilder(**scene)
    sx = -(n - 1) * spacing * 0.5
    sz = -(m - 1) * spacing * 0.5
    parts = np.zeros(n * m, np.int64)
    for j in range(m):
        for i in range(n):
            parts[i + j * n] = b.add_particle(mass, 0.5, (sx + i * spacing, y, sz + j * spacing),
                                              is_static=(j == 0 and (i == 0 or i == n - 1)))
    for j in range(m - 1):
        for i in range(n - 1):
            v00, v10 = parts[i + j * n], parts[i + 1 + j * n]
            v01, v11 = parts[i + (j + 1) * n], parts[i + 1 + (j + 1) * n]
            b.add_tri(v00, v10, v11, E, nu)
            b.add_tri(v00, v11, v01, E, nu)
    return b.build()


def random_quaternion(rng):
    """Uniform random unit quaternion (Shoemake 1992), returned (w, x, y, z)."""
    u1, u2, u3 = rng.random(3)
    a, c = math.sqrt(1 - u1), math.sqrt(u1)
    x = a * math.sin(2 * math.pi * u2)
    y = a * math.cos(2 * math.pi * u2)
    z = c * math.sin(2 * math.pi * u3)
    w = c * math.cos(2 * math.pi * u3)
    q = np.array([w, x, y, z], np.float64)
    return tuple(float(v) for v in (q / np.linalg.norm(q)).astype(F32))


def box_lattice(shape=(40, 25, 20), spacing=3.5, jitter=0.25, edge=(0.5, 2.0), seed=20260101, ground_y=-10.0,
                n_chain=0, n_springs=0, spring_k=1e4, bounds=200.0, chain_density=1000.0, **scene):
    """Jittered lattice of randomly sized and oriented boxes above a large ground.

    Optional: a ball-joint chain of `n_chain` links (0.1 m cubes of density
    chain_density) lying on the ground beside the lattice, anchored to the
    world at one end (Joint stiffness (inf, inf, inf, 0, 0, 0)), and `n_springs` springs between
    seeded random lattice-neighbour box pairs."""
    rng = np.random.default_rng(seed)
    params = dict(bounds_min=(-bounds,) * 3, bounds_max=(bounds,) * 3)
    params.update(scene)
    b = WorldBuilder(**params)
    nx, ny, nz = shape
    gx = nx * spacing + 40.0
    gz = nz * spacing + 40.0
    b.add_body(MODEL_CUBE, 1.0, 0.5, (0.0, ground_y, 0.0), (gx, 1.0, gz), is_static=True)
    base_y = ground_y + 0.5 + 2.0
    lattice = np.zeros((nx, ny, nz), np.int64)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                pos = ((i - (nx - 1) * 0.5) * spacing + rng.uniform(-jitter, jitter),
                       base_y + j * spacing + rng.uniform(-jitter, jitter),
                       (k - (nz - 1) * 0.5) * spacing + rng.uniform(-jitter, jitter))
                scale = tuple(rng.uniform(edge[0], edge[1], 3))
                lattice[i, j, k] = b.add_body(MODEL_CUBE, 1.0, 0.5, pos, scale, rotation=random_quaternion(rng))
    if n_chain:
        # The chain lies straight along x on the ground 10 m beside the
        # lattice, its first link anchored to the world; links are 0.1 m
        # cubes of density chain_density (1 kg at the default, comparable to
        # the boxes that tumble into it), 2 mm above the ground top.  Hanging
        # or released from a height, a chain of 1000 hard ball joints never
        # converges in a step's iterations, stretches and within a few
        # seconds spins its links up until the explosion check stops the
        # scene; 1 g links hit by kilogram boxes do the same.  The
        # reference's algorithm (the oracle) behaves alike in both cases.
        link = 0.1
        y = ground_y + 0.5 + link * 0.5 + 0.002
        start = np.array([-(n_chain * link) * 0.5, y, -(nz * spacing) * 0.5 - 10.0])
        ball = (INF, INF, INF, F32(0), F32(0), F32(0))
        last = None
        for c in range(n_chain):
            p = start + np.array([(c + 0.5) * link, 0.0, 0.0])
            body = b.add_body(MODEL_CUBE, chain_density, 0.5, tuple(p), (link, link, link))
            if last is None:
                b.add_joint(None, tuple(start), body, (-0.5, 0.0, 0.0), ball)
            else:
                b.add_joint(last, (0.5, 0.0, 0.0), body, (-0.5, 0.0, 0.0), ball)
            last = body
    if n_springs:
        dirs = np.array([(1, 0, 0), (0, 1, 0), (0, 0, 1)])
        made = 0
        while made < n_springs:
            i, j, k = rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)
            d = dirs[rng.integers(0, 3)]
            i2, j2, k2 = i + d[0], j + d[1], k + d[2]
            if i2 >= nx or j2 >= ny or k2 >= nz:
                continue
            b.add_spring(int(lattice[i, j, k]), (0.0, 0.0, 0.0), int(lattice[i2, j2, k2]), (0.0, 0.0, 0.0), spring_k)
            made += 1
    return b.build()


def sphere_pile(n=24, seed=5, boxes=4, **scene):
    """Spheres of seeded radii (0.3-0.6 m) dropped on a lattice over the
    ground among a few boxes, with sphere collision switched on (scene option
    SCENE_SPHERE_COLLISION: sphere-sphere, sphere-box and sphere-ground
    contacts).  Without the option the same bodies collide as boxes, as in
    the reference (manifold.cpp:57)."""
    scene.setdefault("options", SCENE_SPHERE_COLLISION)
    rng = np.random.default_rng(seed)
    b = WorldBuilder(**scene)
    add_ground(b)
    total = n + boxes
    cols = max(2, int(math.ceil(math.sqrt(total / 2.0))))
    box_slots = set(range(0, total, max(1, total // boxes))[:boxes]) if boxes > 0 else set()
    for i in range(total):
        layer, r = divmod(i, cols * cols)
        ix, iz = r % cols, r // cols
        pos = ((ix - (cols - 1) * 0.5) * 1.4 + rng.uniform(-0.05, 0.05), -9.5 + 0.7 + layer * 1.4,
               (iz - (cols - 1) * 0.5) * 1.4 + rng.uniform(-0.05, 0.05))
        if i in box_slots:
            b.add_body(MODEL_CUBE, 1.0, 0.5, pos, tuple(rng.uniform(0.5, 1.0, 3)), rotation=tuple(random_quaternion(rng)))
        else:
            d = 2.0 * rng.uniform(0.3, 0.6)
            b.add_body(MODEL_SPHERE, 1.0, 0.5, pos, (d, d, d))
    return b.build()


def sliding_sphere(v0=4.0, radius=0.5, friction=0.5, **scene):
    """A solid sphere launched along the ground with no spin (sphere
    collision on): friction turns sliding into rolling at 5/7 v0."""
    scene.setdefault("options", SCENE_SPHERE_COLLISION)
    b = WorldBuilder(**scene)
    b.add_body(MODEL_CUBE, 1.0, friction, (0.0, -10.0, 0.0), (200.0, 1.0, 200.0), is_static=True)
    d = 2.0 * radius
    b.add_body(MODEL_SPHERE, 1.0, friction, (-40.0, -9.5 + radius - 1e-3, 0.0), (d, d, d), velocity=(v0, 0.0, 0.0))
    return b.build()


def joint_chain(n_links=5, **scene):
    """The JointPlayground pendulum (levels.h:203-233): links joined end to end,
    the first anchored to the world; ball joints (angular stiffness 0 on two axes)."""
    b = WorldBuilder(**scene)
    scale = (2.0, 0.7, 0.7)
    stiff = (INF, INF, INF, INF, F32(0), F32(0))
    anchor = (0.0, 5.0, 0.0)
    pos = [anchor[0] + scale[0] / 2, anchor[1], anchor[2]]
    last = None
    for i in range(n_links):
        body = b.add_body(MODEL_CUBE, 1.0, 0.5, tuple(pos), scale)
        b.add_joint(last, (0.5, 0.0, 0.0) if last is not None else anchor, body, (-0.5, 0.0, 0.0), stiff)
        last = body
        pos[0] += scale[0]
    return b.build()


def spring_bridge(n_links=6, fracture_link=3, **scene):
    """The rope bridge of JointPlayground (levels.h:235-300) with a weight."""
    b = WorldBuilder(**scene)
    start = np.array([-6.0, 5.0, -5.0])
    scale = (2.0, 0.5, 3.0)
    spacing = 1.0
    last = None
    pos = start + np.array([scale[0] / 2, 0.0, 0.0])
    left = start - np.array([spacing, 0.0, 0.0])
    for i in range(n_links):
        body = b.add_body(MODEL_CUBE, 1.0, 0.5, tuple(pos), scale)
        for side in (0, 1):
            zs = 1.0 if side == 0 else -1.0
            if last is not None:
                ra, rb = (0.5, 0.0, zs * 0.5), (-0.5, 0.0, zs * 0.5)
            else:
                ra, rb = tuple(left + np.array([0.0, 0.0, zs * 0.5 * scale[2]])), (-0.5, 0.0, zs * 0.5)
            s = b.add_spring(last, ra, body, rb, 5000.0)
            if i == fracture_link:
                b.spring(s)["fracture"] = 2000.0
        last = body
        pos = pos + np.array([scale[0] + spacing, 0.0, 0.0])
    end = start + np.array([scale[0] * n_links + (n_links - 1) * spacing, 0.0, 0.0])
    right = end + np.array([spacing, 0.0, 0.0])
    for side in (0, 1):
        zs = 1.0 if side == 0 else -1.0
        b.add_spring(None, tuple(right + np.array([0.0, 0.0, zs * 0.5 * scale[2]])), last, (0.5, 0.0, zs * 0.5), INF)
    cx = float((start[0] + end[0]) * 0.5)
    b.add_body(MODEL_CUBE, 1.0, 0.5, (cx, 8.0, float(start[2])), (2.0, 2.0, 2.0))
    return b.build()


def mass_springs(n=4, **scene):
    """MassSprings level (levels.h:160-200): boxes hanging on ropes from world anchors."""
    b = WorldBuilder(**scene)
    start = -((n - 1) * 2.0) * 0.5
    for i in range(n):
        body = b.add_body(MODEL_CUBE, 1.0 * 2.5 ** i, 0.5, (start + i * 2.0, 0.0, 0.0), (1.0, 1.0, 1.0))
        anchor = tuple(np.asarray(b.body(body)["pos"], np.float64))
        b.add_spring(None, anchor, body, (0.0, 0.0, 0.0), 1000.0, 3.0, True)
    return b.build()


def out_of_bounds(**scene):
    """Bodies leaving the world bounds (x in [-6, 6]) in the first step:
    the bounds removal (solver.cpp:520-534) takes a fast box with the spring
    that ties it to a resting box, and the fast half of a small Neo-Hookean
    block with the tets that touch it; the scene's last body, a box resting
    on another, moves into a freed index (swap with last)."""
    params = dict(bounds_min=(-6.0, -20.0, -6.0), bounds_max=(6.0, 20.0, 6.0))
    params.update(scene)
    b = WorldBuilder(**params)
    add_ground(b)
    rest = b.add_body(MODEL_CUBE, 1.0, 0.5, (0.0, -9.0, 0.0), (1.0, 1.0, 1.0))
    fast = b.add_body(MODEL_CUBE, 1.0, 0.5, (5.5, -5.0, 0.0), (1.0, 1.0, 1.0), velocity=(60.0, 0.0, 0.0))
    anchor = b.add_body(MODEL_CUBE, 1.0, 0.5, (-2.0, -9.0, 0.0), (1.0, 1.0, 1.0))
    b.add_spring(anchor, (0.0, 0.0, 0.0), fast, (0.0, 0.0, 0.0), 100.0)
    idx = tet_volume(b, (5.0, -6.0, -0.5), (3, 2, 2), 0.5, 1e4, 0.3, 0.1)
    for p in idx[2].ravel():  # the x = 6.0 column leaves
        b._bodies[int(p)]["vel"] = (60.0, 0.0, 0.0)
    b.add_body(MODEL_CUBE, 1.0, 0.5, (0.0, -7.98, 0.0), (1.0, 1.0, 1.0))
    del rest
    return b.build()


def batched_mix(n_scenes=4096, seed=7, box_range=(8, 2048), tet_range=(1000, 50000), fem_fraction=0.5):
    """Independent scenes of uneven size: rigid box lattices/pyramids with box
    counts log-uniform in box_range, and Neo-Hookean grids with tet counts
    log-uniform in tet_range; each scene has dt in {1/60, 1/120} and 8-16
    iterations."""
    rng = np.random.default_rng(seed)
    worlds = []
    for s in range(n_scenes):
        dt = F32(1.0) / F32(60.0 if rng.random() < 0.5 else 120.0)
        iters = int(rng.integers(8, 17))
        if rng.random() >= fem_fraction:
            nb = int(round(math.exp(rng.uniform(math.log(box_range[0]), math.log(box_range[1])))))
            worlds.append(_random_box_scene(rng, nb, dt=dt, iterations=iters))
        else:
            nt = math.exp(rng.uniform(math.log(tet_range[0]), math.log(tet_range[1])))
            cells = max(1, int(round((nt / 5.0) ** (1.0 / 3.0))))
            nv = cells + 1
            worlds.append(fem_block(n=nv, cell=0.25, E=1e5, nu=0.3, rho=1000.0, dt=dt, iterations=iters))
    return World.concat(worlds)


def _random_box_scene(rng, n_boxes, **scene):
    """A seeded rigid scene with exactly n_boxes dynamic boxes: a lattice of
    randomly sized, randomly oriented boxes above the standard ground."""
    b = WorldBuilder(**scene)
    add_ground(b)
    side = max(1, int(math.ceil(n_boxes ** (1.0 / 3.0))))
    spacing = 2.6
    made = 0
    for j in range(side * 4):
        for k in range(side):
            for i in range(side):
                if made == n_boxes:
                    break
                pos = ((i - (side - 1) * 0.5) * spacing * 0.5, -8.0 + j * spacing, (k - (side - 1) * 0.5) * spacing * 0.5)
                if side * spacing * 0.5 > 9.0:
                    pos = ((i - (side - 1) * 0.5) * spacing * 0.5 * 9.0 / (side * spacing * 0.5), pos[1],
                           (k - (side - 1) * 0.5) * spacing * 0.5 * 9.0 / (side * spacing * 0.5))
                scale = tuple(rng.uniform(0.5, 1.2, 3))
                b.add_body(MODEL_CUBE, 1.0, 0.5, pos, scale, rotation=random_quaternion(rng))
                made += 1
    return b.build()
