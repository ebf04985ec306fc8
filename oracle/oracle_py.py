"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Only the tests, the smoke check and the benchmark's CPU-baseline leg load
this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from skyephyx_amd.abi import AvbdWorld, world_to_c, alloc_world
from skyephyx_amd.world import World

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

NATURAL, COLOUR_MAJOR = 0, 1


def build(force=False):
    """Compile the oracle with its Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_load.argtypes = [C.c_void_p, C.POINTER(AvbdWorld), C.c_int]
        L.orc_set_order.argtypes = [C.c_void_p, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_remove_bodies.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int]
        L.orc_set_drag.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                   C.c_int]
        L.orc_debug.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.POINTER(C.c_int64),
                                C.POINTER(C.c_int64)]
        L.orc_counts.argtypes = [C.c_void_p, C.POINTER(AvbdWorld)]
        L.orc_save.argtypes = [C.c_void_p, C.POINTER(AvbdWorld)]
        L.orc_colours.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int]
        L.orc_timings.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        L.orc_collide.argtypes = [fp, fp, C.c_int, fp, C.POINTER(C.c_uint32)]
        L.orc_collide_bodies.argtypes = [fp, fp, C.c_int, C.c_int, C.c_int, fp, C.POINTER(C.c_uint32)]
        L.orc_nh_energy_change.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_int, dp, C.c_int]
        L.orc_nh_energy_change.restype = C.c_double
        L.orc_stvk_energy_change.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_int, dp, C.c_int]
        L.orc_stvk_energy_change.restype = C.c_double
        L.orc_ldlt_solve.argtypes = [dp, dp, dp, C.c_int]
        L.orc_nh_terms.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_stvk_terms.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_svd3.argtypes = [dp, C.c_int, dp, dp, dp]
        L.orc_sym_eig3.argtypes = [dp, dp, dp]
        L.orc_rot_helper.argtypes = [C.c_int, fp, fp, fp]
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Oracle:
    """One scene of a world, stepped on the CPU by the restatement."""

    def __init__(self, world: World, scene: int = 0, fp64: bool = False, order: int = NATURAL):
        self.L = lib()
        self.h = self.L.orc_create(1 if fp64 else 0)
        keep, cw = world_to_c(world)
        rc = self.L.orc_load(self.h, C.byref(cw), scene)
        if rc != 0:
            raise RuntimeError(f"orc_load failed ({rc})")
        self.L.orc_set_order(self.h, order)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_order(self, order):
        self.L.orc_set_order(self.h, order)

    def remove_bodies(self, idx):
        """Solver::RemoveBody of the bodies at these indices (before the call), in order."""
        a = np.ascontiguousarray(idx, np.int32)
        if self.L.orc_remove_bodies(self.h, a.ctypes.data_as(C.POINTER(C.c_int32)), len(a)) != 0:
            raise ValueError("body index out of range")

    def set_drag(self, idx, added_velocity, dragged=True):
        """Mesh::isDragged / addedDragVelocity (geometry.hpp:193-194) of the
        bodies at these indices; the next velocity update adds the velocity
        once (solver.cpp:506-510)."""
        i = np.ascontiguousarray(idx, np.int32)
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(added_velocity, np.float32), (len(i), 3)))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dragged, np.int32), (len(i),)))
        if self.L.orc_set_drag(self.h, i.ctypes.data_as(C.POINTER(C.c_int32)), _fptr(v),
                               d.ctypes.data_as(C.POINTER(C.c_int32)), len(i)) != 0:
            raise ValueError("body index out of range")

    def debug(self):
        """(lines [n, 10], points [n, 8]) of the last step (the renderer's data)."""
        nl, npt = C.c_int64(), C.c_int64()
        self.L.orc_debug(self.h, None, 0, None, 0, C.byref(nl), C.byref(npt))
        lines = np.zeros((nl.value, 10), np.float32)
        points = np.zeros((npt.value, 8), np.float32)
        self.L.orc_debug(self.h, lines.ctypes.data, nl.value, points.ctypes.data, npt.value, C.byref(nl),
                         C.byref(npt))
        return lines, points

    def step(self, n=1):
        self.L.orc_step(self.h, n)

    def save(self) -> World:
        cnt = AvbdWorld()
        self.L.orc_counts(self.h, C.byref(cnt))
        w, cw, keep = alloc_world(cnt)
        rc = self.L.orc_save(self.h, C.byref(cw))
        if rc != 0:
            raise RuntimeError(f"orc_save failed ({rc})")
        return w

    def colours(self, n):
        out = np.full(n, -2, np.int32)
        nc = self.L.orc_colours(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), n)
        return out, nc

    def timings(self):
        t = np.zeros(7, np.float64)
        self.L.orc_timings(self.h, _dptr(t))
        return dict(zip(["broad", "warm", "predict", "primal_dual", "velocity", "post_stab", "total"], t))


def collide(box_a, box_b, fp64=False):
    """box = (pos[3], rot[4] (w,x,y,z), scale[3]) -> (n, contacts (n,13), ids (n,))."""
    a = np.asarray(box_a, np.float32).reshape(10)
    b = np.asarray(box_b, np.float32).reshape(10)
    out = np.zeros((8, 13), np.float32)
    ids = np.zeros(8, np.uint32)
    n = lib().orc_collide(_fptr(a), _fptr(b), 1 if fp64 else 0, _fptr(out), ids.ctypes.data_as(C.POINTER(C.c_uint32)))
    return n, out[:n].copy(), ids[:n].copy()


def collide_bodies(body_a, body_b, sphere_a=False, sphere_b=False, fp64=False):
    """Contacts of two bodies (pos[3], rot[4], scale[3]), each possibly a
    sphere of radius scale[0] / 2 -> (n, contacts (n,13), ids (n,))."""
    a = np.asarray(body_a, np.float32).reshape(10)
    b = np.asarray(body_b, np.float32).reshape(10)
    out = np.zeros((8, 13), np.float32)
    ids = np.zeros(8, np.uint32)
    n = lib().orc_collide_bodies(_fptr(a), _fptr(b), int(sphere_a), int(sphere_b), 1 if fp64 else 0, _fptr(out),
                                 ids.ctypes.data_as(C.POINTER(C.c_uint32)))
    return n, out[:n].copy(), ids[:n].copy()


def neohookean_energy_change(p, dm_inv, V0, mu, lam, vtx, dx, fp64=True):
    """E(vertex vtx at p[vtx] - dx) - E(p) of one Neo-Hookean tet."""
    p = np.ascontiguousarray(p, np.float64).reshape(12)
    dm = np.ascontiguousarray(dm_inv, np.float64).reshape(9)
    d = np.ascontiguousarray(dx, np.float64).reshape(3)
    return lib().orc_nh_energy_change(_dptr(p), _dptr(dm), V0, mu, lam, vtx, _dptr(d), 1 if fp64 else 0)


def stvk_energy_change(p, dm_inv, area, mu, lam, vtx, dx, fp64=True):
    """E(vertex vtx at p[vtx] - dx) - E(p) of one StVK triangle."""
    p = np.ascontiguousarray(p, np.float64).reshape(9)
    dm = np.ascontiguousarray(dm_inv, np.float64).reshape(4)
    d = np.ascontiguousarray(dx, np.float64).reshape(3)
    return lib().orc_stvk_energy_change(_dptr(p), _dptr(dm), area, mu, lam, vtx, _dptr(d), 1 if fp64 else 0)


def ldlt_solve(A, b, fp64=True):
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(6, np.float64)
    lib().orc_ldlt_solve(_dptr(A), _dptr(b), _dptr(x), 1 if fp64 else 0)
    return x


def neohookean_terms(p, dm_inv, V0, mu, lam, vtx, proj=1, fp64=True):
    p = np.ascontiguousarray(p, np.float64).reshape(12)
    dm = np.ascontiguousarray(dm_inv, np.float64).reshape(9)
    g = np.zeros(3)
    H = np.zeros(9)
    lib().orc_nh_terms(_dptr(p), _dptr(dm), V0, mu, lam, vtx, proj, 1 if fp64 else 0, _dptr(g), _dptr(H))
    return g, H.reshape(3, 3)


def stvk_terms(p, dm_inv, area, mu, lam, vtx, proj=1, fp64=True):
    p = np.ascontiguousarray(p, np.float64).reshape(9)
    dm = np.ascontiguousarray(dm_inv, np.float64).reshape(4)
    g = np.zeros(3)
    H = np.zeros(9)
    lib().orc_stvk_terms(_dptr(p), _dptr(dm), area, mu, lam, vtx, proj, 1 if fp64 else 0, _dptr(g), _dptr(H))
    return g, H.reshape(3, 3)


def svd3(F, fp64=True):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    U = np.zeros(9)
    S = np.zeros(3)
    V = np.zeros(9)
    lib().orc_svd3(_dptr(F), 1 if fp64 else 0, _dptr(U), _dptr(S), _dptr(V))
    return U.reshape(3, 3), S, V.reshape(3, 3)


def sym_eig3(A):
    A = np.ascontiguousarray(A, np.float64).reshape(9)
    w = np.zeros(3)
    V = np.zeros(9)
    lib().orc_sym_eig3(_dptr(A), _dptr(w), _dptr(V))
    return w, V.reshape(3, 3)


def rot_helper(op, a, b=None):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b if b is not None else np.zeros(4), np.float32)
    out = np.zeros(9, np.float32)
    n = lib().orc_rot_helper(op, _fptr(a), _fptr(b), _fptr(out))
    return out[:n].copy()
