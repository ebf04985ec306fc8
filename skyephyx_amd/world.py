"""World description: numpy record arrays that mirror the C ABI records in
include/avbd_capi.h, and a builder with the reference's construction API.

The builder restates what the reference's constructors compute when a scene
is built (src/physics/solver.cpp:50-179 AddBody/AddParticle, joint.cpp:24-66,
spring.cpp:13-38, neohookeanfem.cpp:5-37, stvkfem.cpp:5-35) so that a world
can be created without the C++ facade: bodies get mass, diagonal inertia and
detection radius, joints their rest rotation and torque arm, springs their
rest length, tets and triangles their rest-shape inverse, volume/area and
Lame parameters.  FP32 arithmetic is used where the reference's constructor
result feeds the dynamics directly (masses, inertia); rest-shape inverses are
formed in FP64 and rounded once.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32

BODY_STATIC = 1
BODY_PARTICLE = 2
BODY_SPHERE = 4
SCENE_TRUST_REGION = 1
SCENE_SPHERE_COLLISION = 2
FORCE_DISABLED = 1
PROJ_CLAMP, PROJ_ABSOLUTE, PROJ_ADAPTIVE = 0, 1, 2

MODEL_CUBE, MODEL_SPHERE, MODEL_PYRAMID = 0, 1, 2

SCENE_DTYPE = np.dtype(
    [
        ("dt", "<f4"),
        ("iterations", "<i4"),
        ("alpha", "<f4"),
        ("beta", "<f4"),
        ("gamma", "<f4"),
        ("on_penetration_penalty", "<f4"),
        ("post_stabilization", "<i4"),
        ("projection_mode", "<i4"),
        ("gravity", "<f4", (3,)),
        ("bounds_min", "<f4", (3,)),
        ("bounds_max", "<f4", (3,)),
        ("emergency_stop", "<i4"),
        ("n_removed", "<i4"),
        ("options", "<u4"),
        ("trust_region_rho", "<f4"),
    ],
    align=True,
)

BODY_DTYPE = np.dtype(
    [
        ("pos", "<f4", (3,)),
        ("rot", "<f4", (4,)),
        ("scale", "<f4", (3,)),
        ("vel", "<f4", (3,)),
        ("ang_vel", "<f4", (3,)),
        ("prev_vel", "<f4", (3,)),
        ("prev_ang_vel", "<f4", (3,)),
        ("mass", "<f4"),
        ("inertia", "<f4", (3,)),
        ("friction", "<f4"),
        ("radius", "<f4"),
        ("flags", "<u4"),
        ("scene", "<i4"),
    ],
    align=True,
)

JOINT_DTYPE = np.dtype(
    [
        ("body_a", "<i4"),
        ("body_b", "<i4"),
        ("serial", "<i8"),
        ("flags", "<u4"),
        ("r_a", "<f4", (3,)),
        ("r_b", "<f4", (3,)),
        ("rest_rot", "<f4", (4,)),
        ("torque_arm", "<f4"),
        ("stiffness", "<f4", (6,)),
        ("fracture", "<f4", (6,)),
        ("fmin", "<f4", (6,)),
        ("fmax", "<f4", (6,)),
        ("lambda", "<f4", (6,)),
        ("penalty", "<f4", (6,)),
    ],
    align=True,
)

SPRING_DTYPE = np.dtype(
    [
        ("body_a", "<i4"),
        ("body_b", "<i4"),
        ("serial", "<i8"),
        ("flags", "<u4"),
        ("r_a", "<f4", (3,)),
        ("r_b", "<f4", (3,)),
        ("rest_length", "<f4"),
        ("stiffness", "<f4"),
        ("fracture", "<f4"),
        ("fmin", "<f4"),
        ("fmax", "<f4"),
        ("lambda", "<f4"),
        ("penalty", "<f4"),
    ],
    align=True,
)

CONTACT_DTYPE = np.dtype(
    [
        ("id", "<u4"),
        ("stick", "<i4"),
        ("r_a", "<f4", (3,)),
        ("r_b", "<f4", (3,)),
        ("normal", "<f4", (3,)),
        ("penetration", "<f4"),
        ("lambda", "<f4", (3,)),
        ("penalty", "<f4", (3,)),
    ],
    align=True,
)

MANIFOLD_DTYPE = np.dtype(
    [
        ("body_a", "<i4"),
        ("body_b", "<i4"),
        ("serial", "<i8"),
        ("n_contacts", "<i4"),
        ("friction", "<f4"),
        ("contacts", CONTACT_DTYPE, (8,)),
    ],
    align=True,
)

TET_DTYPE = np.dtype(
    [
        ("v", "<i4", (4,)),
        ("serial", "<i4"),
        ("dm_inv", "<f4", (9,)),
        ("rest_volume", "<f4"),
        ("mu", "<f4"),
        ("lambda", "<f4"),
    ],
    align=True,
)

TRI_DTYPE = np.dtype(
    [
        ("v", "<i4", (3,)),
        ("serial", "<i4"),
        ("dm_inv", "<f4", (4,)),
        ("rest_area", "<f4"),
        ("mu", "<f4"),
        ("lambda", "<f4"),
    ],
    align=True,
)

RECORD_DTYPES = [SCENE_DTYPE, BODY_DTYPE, JOINT_DTYPE, SPRING_DTYPE, CONTACT_DTYPE, MANIFOLD_DTYPE, TET_DTYPE, TRI_DTYPE]

INF = F32(np.inf)


def default_scene(**kw) -> np.ndarray:
    """Solver defaults (solver.hpp:69-84), gravity solver.cpp:7, bounds constants.hpp:11-12."""
    s = np.zeros((), dtype=SCENE_DTYPE)
    s["dt"] = F32(1.0) / F32(60.0)
    s["iterations"] = 10
    s["alpha"] = 0.95
    s["beta"] = 100000.0
    s["gamma"] = 0.99
    s["on_penetration_penalty"] = 0.0
    s["post_stabilization"] = 1
    s["projection_mode"] = PROJ_ABSOLUTE
    s["gravity"] = (0.0, -9.81, 0.0)
    s["bounds_min"] = (-50.0, -50.0, -50.0)
    s["bounds_max"] = (50.0, 50.0, 50.0)
    for k, v in kw.items():
        s[k] = v
    return s


@dataclass
class World:
    """A world of one or more scenes in ABI record form."""

    scenes: np.ndarray
    bodies: np.ndarray
    joints: np.ndarray
    springs: np.ndarray
    manifolds: np.ndarray
    tets: np.ndarray
    tris: np.ndarray

    @staticmethod
    def empty(n_scenes=0) -> "World":
        return World(
            np.zeros(n_scenes, SCENE_DTYPE),
            np.zeros(0, BODY_DTYPE),
            np.zeros(0, JOINT_DTYPE),
            np.zeros(0, SPRING_DTYPE),
            np.zeros(0, MANIFOLD_DTYPE),
            np.zeros(0, TET_DTYPE),
            np.zeros(0, TRI_DTYPE),
        )

    def copy(self) -> "World":
        return World(*(np.array(a, copy=True) for a in self.arrays()))

    def arrays(self):
        return (self.scenes, self.bodies, self.joints, self.springs, self.manifolds, self.tets, self.tris)

    def scene_bodies(self, s: int) -> np.ndarray:
        return np.nonzero(self.bodies["scene"] == s)[0]

    def extract_scene(self, s: int) -> "World":
        """One scene as a single-scene world (body indices renumbered)."""
        idx = self.scene_bodies(s)
        remap = -np.ones(len(self.bodies) + 1, np.int64)
        remap[idx] = np.arange(len(idx))
        b = self.bodies[idx].copy()
        b["scene"] = 0

        def take(arr, keys):
            if len(arr) == 0:
                return arr.copy()
            first = arr[keys[-1]] if arr[keys[-1]].ndim == 1 else arr[keys[-1]][:, 0]
            sel = np.isin(first, idx)
            out = arr[sel].copy()
            for k in keys:
                v = out[k]
                out[k] = np.where(v >= 0, remap[np.maximum(v, 0)], -1)
            return out

        return World(
            self.scenes[s : s + 1].copy(),
            b,
            take(self.joints, ["body_a", "body_b"]),
            take(self.springs, ["body_a", "body_b"]),
            take(self.manifolds, ["body_a", "body_b"]),
            take(self.tets, ["v"]),
            take(self.tris, ["v"]),
        )

    @staticmethod
    def concat(worlds) -> "World":
        """Concatenate single- or multi-scene worlds into one world."""
        out = World.empty(0)
        parts = {k: [] for k in ("scenes", "bodies", "joints", "springs", "manifolds", "tets", "tris")}
        body_off = 0
        scene_off = 0
        serial_off = 0
        for w in worlds:
            parts["scenes"].append(w.scenes)
            b = w.bodies.copy()
            b["scene"] += scene_off
            parts["bodies"].append(b)
            for name in ("joints", "springs", "manifolds"):
                a = getattr(w, name).copy()
                if len(a):
                    a["body_a"] = np.where(a["body_a"] >= 0, a["body_a"] + body_off, -1)
                    a["body_b"] = a["body_b"] + body_off
                    a["serial"] += serial_off
                parts[name].append(a)
            for name in ("tets", "tris"):
                a = getattr(w, name).copy()
                if len(a):
                    a["v"] += body_off
                parts[name].append(a)
            mx = -1
            for name in ("joints", "springs", "manifolds"):
                a = getattr(w, name)
                if len(a):
                    mx = max(mx, int(a["serial"].max()))
            serial_off += mx + 1
            body_off += len(w.bodies)
            scene_off += len(w.scenes)
        for k, v in parts.items():
            if v:
                setattr(out, k, np.concatenate(v))
        return out


def _quat_mul(a, b):
    aw, ax, ay, az = (F32(x) for x in a)
    bw, bx, by, bz = (F32(x) for x in b)
    return np.array(
        [
            ((aw * bw - ax * bx) - ay * by) - az * bz,
            ((aw * bx + ax * bw) + ay * bz) - az * by,
            ((aw * by - ax * bz) + ay * bw) + az * bx,
            ((aw * bz + ax * by) - ay * bx) + az * bw,
        ],
        F32,
    )


def _rotmat(q):
    w, x, y, z = (F32(v) for v in q)
    two = F32(2)
    tx, ty, tz = two * x, two * y, two * z
    twx, twy, twz = tx * w, ty * w, tz * w
    txx, txy, txz = tx * x, ty * x, tz * x
    tyy, tyz, tzz = ty * y, tz * y, tz * z
    one = F32(1)
    return np.array(
        [
            [one - (tyy + tzz), txy - twz, txz + twy],
            [txy + twz, one - (txx + tzz), tyz - twx],
            [txz - twy, tyz + twx, one - (txx + tyy)],
        ],
        F32,
    )


def _dot3(a, b):
    return (F32(a[0]) * F32(b[0]) + F32(a[1]) * F32(b[1])) + F32(a[2]) * F32(b[2])


def transform_point(body, p):
    """Transform::TransformPoint (geometry.hpp:64-67): R(q) (s . p) + x."""
    R = _rotmat(body["rot"])
    sp = np.asarray(body["scale"], F32) * np.asarray(p, F32)
    return np.array([_dot3(R[i], sp) for i in range(3)], F32) + np.asarray(body["pos"], F32)


class WorldBuilder:
    """Builds one scene with the reference's construction API.

    Method names and argument meanings follow Solver (solver.hpp:43-46),
    Joint/Spring (force.hpp:95,119-122) and NeoHookeanFEM/STVKFEM
    (energy.hpp:50-52,90-92); bodies are returned as integer indices.
    """

    def __init__(self, **scene_params):
        self.scene = default_scene(**scene_params)
        self._bodies = []
        self._joints = []
        self._springs = []
        self._tets = []
        self._tris = []
        self._force_serial = 0
        self._energy_serial = 0

    # ---- bodies -------------------------------------------------------
    def add_body(self, model_type, density, friction, position, scale, velocity=(0, 0, 0), rotation=(1, 0, 0, 0),
                 angular_velocity=(0, 0, 0), is_static=False):
        """Solver::AddBody (solver.cpp:50-101)."""
        if model_type == MODEL_PYRAMID and not is_static:
            raise ValueError("pyramid bodies leave the inertia undefined in the reference (solver.cpp:77-90)")
        b = np.zeros((), BODY_DTYPE)
        sc = np.asarray(scale, F32)
        b["pos"] = position
        b["rot"] = rotation
        b["scale"] = sc
        b["vel"] = velocity
        b["ang_vel"] = angular_velocity
        b["friction"] = friction
        b["radius"] = np.sqrt(_dot3(sc, sc)).astype(F32) * F32(0.5)
        density = F32(density)
        if is_static:
            b["flags"] = BODY_STATIC
            b["mass"] = 0.0
            b["inertia"] = 0.0
        else:
            sx, sy, sz = sc
            sx2, sy2, sz2 = sx * sx, sy * sy, sz * sz
            if model_type == MODEL_SPHERE:
                r = sx * F32(0.5)
                m = F32(float(density * (F32(4.0) / F32(3.0))) * math.pi * float(r) * float(r) * float(r))
                inertia = np.full(3, (F32(2.0) / F32(5.0)) * m * r * r, F32)
            else:
                m = ((density * sx) * sy) * sz
                k = m / F32(12.0)
                inertia = np.array([k * (sy2 + sz2), k * (sx2 + sz2), k * (sx2 + sy2)], F32)
            b["mass"] = m
            b["inertia"] = inertia
        if model_type == MODEL_SPHERE:
            b["flags"] = b["flags"] | BODY_SPHERE
        self._bodies.append(b)
        return len(self._bodies) - 1

    def add_particle(self, mass, friction, position, velocity=(0, 0, 0), is_static=False):
        """Solver::AddParticle (solver.cpp:119-131): a 0.1 m cube with zero inertia."""
        s = F32(0.1)
        density = F32(mass) / ((s * s) * s)
        i = self.add_body(MODEL_CUBE, density, friction, position, (s, s, s), velocity, (1, 0, 0, 0), (0, 0, 0),
                          is_static)
        b = self._bodies[i]
        b["flags"] = int(b["flags"]) | BODY_PARTICLE
        b["inertia"] = 0.0
        b["ang_vel"] = 0.0
        return i

    def body(self, i):
        return self._bodies[i]

    @property
    def n_bodies(self):
        return len(self._bodies)

    # ---- forces -------------------------------------------------------
    def _next_force_serial(self):
        s = self._force_serial
        self._force_serial += 1
        return s

    def add_joint(self, body_a, r_a, body_b, r_b, stiffness=(INF,) * 6, fractures=(INF,) * 6):
        """Joint ctor (joint.cpp:24-66). body_a may be None (world anchor)."""
        j = np.zeros((), JOINT_DTYPE)
        j["body_a"] = -1 if body_a is None else body_a
        j["body_b"] = body_b
        j["serial"] = self._next_force_serial()
        j["r_a"] = r_a
        j["r_b"] = r_b
        j["stiffness"] = np.asarray(stiffness, F32)
        fr = np.asarray(fractures, F32)
        j["fracture"] = fr
        j["fmin"] = -fr
        j["fmax"] = fr
        qa = (1, 0, 0, 0) if body_a is None else self._bodies[body_a]["rot"]
        qb = self._bodies[body_b]["rot"]
        qa = np.asarray(qa, F32)
        j["rest_rot"] = _quat_mul((qa[0], -qa[1], -qa[2], -qa[3]), qb)
        sa = np.zeros(3, F32) if body_a is None else np.asarray(self._bodies[body_a]["scale"], F32)
        ssum = sa + np.asarray(self._bodies[body_b]["scale"], F32)
        arm = _dot3(ssum, ssum)
        j["torque_arm"] = F32(1.0) if arm < F32(1e-6) else arm
        self._joints.append(j)
        return len(self._joints) - 1

    def add_spring(self, body_a, r_a, body_b, r_b, stiffness, rest_length=-1.0, is_rope=False):
        """Spring ctor (spring.cpp:13-38). body_a may be None."""
        s = np.zeros((), SPRING_DTYPE)
        s["body_a"] = -1 if body_a is None else body_a
        s["body_b"] = body_b
        s["serial"] = self._next_force_serial()
        s["r_a"] = r_a
        s["r_b"] = r_b
        s["stiffness"] = stiffness
        s["fracture"] = INF
        s["fmin"] = 0.0 if is_rope else -INF
        s["fmax"] = INF
        if rest_length < 0:
            pa = np.asarray(r_a, F32) if body_a is None else transform_point(self._bodies[body_a], r_a)
            pb = transform_point(self._bodies[body_b], r_b)
            d = pb - pa
            rest_length = np.sqrt(_dot3(d, d)).astype(F32)
        s["rest_length"] = rest_length
        self._springs.append(s)
        return len(self._springs) - 1

    def spring(self, i):
        return self._springs[i]

    # ---- energies -----------------------------------------------------
    @staticmethod
    def lame(E, nu):
        E, nu = F32(E), F32(nu)
        mu = E / (F32(2.0) * (F32(1.0) + nu))
        lam = E * nu / ((F32(1.0) + nu) * (F32(1.0) - F32(2.0) * nu))
        return mu, lam

    def add_tet(self, v0, v1, v2, v3, E, nu):
        """NeoHookeanFEM ctor (neohookeanfem.cpp:5-37)."""
        t = np.zeros((), TET_DTYPE)
        t["v"] = (v0, v1, v2, v3)
        t["serial"] = self._energy_serial
        self._energy_serial += 1
        p = [np.asarray(self._bodies[i]["pos"], np.float64) for i in (v0, v1, v2, v3)]
        Dm = np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1)
        t["dm_inv"] = np.linalg.inv(Dm).astype(F32).reshape(-1)
        t["rest_volume"] = F32(abs(np.linalg.det(Dm)) / 6.0)
        t["mu"], t["lambda"] = self.lame(E, nu)
        self._tets.append(t)
        return len(self._tets) - 1

    def add_tri(self, v0, v1, v2, E, nu):
        """STVKFEM ctor (stvkfem.cpp:5-35): 2-D rest frame of the triangle."""
        t = np.zeros((), TRI_DTYPE)
        t["v"] = (v0, v1, v2)
        t["serial"] = self._energy_serial
        self._energy_serial += 1
        p = [np.asarray(self._bodies[i]["pos"], np.float64) for i in (v0, v1, v2)]
        e0, e1 = p[1] - p[0], p[2] - p[0]
        x1 = np.linalg.norm(e0)
        x2 = e1.dot(e0) / x1
        y2 = math.sqrt(max(0.0, e1.dot(e1) - x2 * x2))
        Dm = np.array([[x1, x2], [0.0, y2]])
        t["dm_inv"] = np.linalg.inv(Dm).astype(F32).reshape(-1)
        t["rest_area"] = F32(0.5 * x1 * y2)
        t["mu"], t["lambda"] = self.lame(E, nu)
        self._tris.append(t)
        return len(self._tris) - 1

    # ---- output -------------------------------------------------------
    def build(self) -> World:
        def arr(lst, dt):
            return np.array(lst, dtype=dt) if lst else np.zeros(0, dt)

        return World(
            np.array([self.scene], SCENE_DTYPE),
            arr(self._bodies, BODY_DTYPE),
            arr(self._joints, JOINT_DTYPE),
            arr(self._springs, SPRING_DTYPE),
            np.zeros(0, MANIFOLD_DTYPE),
            arr(self._tets, TET_DTYPE),
            arr(self._tris, TRI_DTYPE),
        )
