"""ctypes view of the avbd_world exchange struct (include/avbd_capi.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .world import (
    BODY_DTYPE,
    JOINT_DTYPE,
    MANIFOLD_DTYPE,
    SCENE_DTYPE,
    SPRING_DTYPE,
    TET_DTYPE,
    TRI_DTYPE,
    World,
)


class AvbdWorld(C.Structure):
    _fields_ = [
        ("n_scenes", C.c_int32),
        ("n_bodies", C.c_int32),
        ("n_joints", C.c_int32),
        ("n_springs", C.c_int32),
        ("n_manifolds", C.c_int32),
        ("n_tets", C.c_int32),
        ("n_tris", C.c_int32),
        ("scenes", C.c_void_p),
        ("bodies", C.c_void_p),
        ("joints", C.c_void_p),
        ("springs", C.c_void_p),
        ("manifolds", C.c_void_p),
        ("tets", C.c_void_p),
        ("tris", C.c_void_p),
    ]


class AvbdTimings(C.Structure):
    _fields_ = [
        ("broad_phase_ms", C.c_float),
        ("warmstart_ms", C.c_float),
        ("prediction_ms", C.c_float),
        ("primal_dual_ms", C.c_float),
        ("velocity_update_ms", C.c_float),
        ("post_stab_ms", C.c_float),
        ("total_ms", C.c_float),
        ("n_colours", C.c_int32),
    ]


class AvbdRequestInfo(C.Structure):
    _fields_ = [
        ("ticket", C.c_int64),
        ("scene", C.c_int32),
        ("steps", C.c_int32),
        ("state", C.c_int32),
        ("pad_", C.c_int32),
        ("t_arrival", C.c_double),
        ("t_start", C.c_double),
        ("t_done", C.c_double),
    ]


class AvbdServerStats(C.Structure):
    _fields_ = [
        ("ticks", C.c_int64),
        ("scene_steps", C.c_int64),
        ("completed", C.c_int64),
        ("layout_rebuilds", C.c_int64),
        ("resident_bytes", C.c_int64),
        ("resident_scenes", C.c_int32),
        ("queued", C.c_int32),
        ("running", C.c_int32),
        ("pad_", C.c_int32),
    ]


_NAMES = ("scenes", "bodies", "joints", "springs", "manifolds", "tets", "tris")
_DTYPES = (SCENE_DTYPE, BODY_DTYPE, JOINT_DTYPE, SPRING_DTYPE, MANIFOLD_DTYPE, TET_DTYPE, TRI_DTYPE)


def _ptr(a):
    return a.ctypes.data if a.size else None


def world_to_c(w: World):
    """Build an AvbdWorld pointing at contiguous copies of the world's arrays.
    Returns (keepalive, struct)."""
    arrs = [np.ascontiguousarray(getattr(w, n), dtype=dt) for n, dt in zip(_NAMES, _DTYPES)]
    cw = AvbdWorld()
    for n, a in zip(_NAMES, arrs):
        setattr(cw, "n_" + n, len(a))
        setattr(cw, n, _ptr(a))
    return arrs, cw


def alloc_world(counts: AvbdWorld):
    """Allocate zeroed arrays for the counts in `counts`; returns (World, struct, keepalive)."""
    arrs = [np.zeros(getattr(counts, "n_" + n), dt) for n, dt in zip(_NAMES, _DTYPES)]
    w = World(*arrs)
    cw = AvbdWorld()
    for n, a in zip(_NAMES, arrs):
        setattr(cw, "n_" + n, len(a))
        setattr(cw, n, _ptr(a))
    return w, cw, arrs
