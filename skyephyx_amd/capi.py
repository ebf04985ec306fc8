"""ctypes binding of libavbd_hip.so (the C ABI in include/avbd_capi.h).

There is no fallback: if the HIP library is missing or no gfx950 device is
present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import AvbdRequestInfo, AvbdServerStats, AvbdTimings, AvbdWorld, alloc_world, world_to_c
from .world import BODY_DTYPE, World

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libavbd_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "avbd_capi.h")

_lib = None


NOALLOC_MAX_BODIES = 131072  # AVBD_NOALLOC_MAX_BODIES


class AvbdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"avbd error {code}: {msg}")
        self.code = code


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build the HIP extension first (__graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.avbd_abi_version.restype = C.c_int
        L.avbd_record_sizes.argtypes = [C.POINTER(C.c_int32), C.c_int32]
        L.avbd_last_error.restype = C.c_char_p
        L.avbd_create.argtypes = [C.c_int32, C.POINTER(vp)]
        L.avbd_destroy.argtypes = [vp]
        L.avbd_upload.argtypes = [vp, C.POINTER(AvbdWorld), C.c_int64]
        L.avbd_step.argtypes = [vp, C.c_int32]
        L.avbd_step_scenes.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32, C.c_int32]
        L.avbd_synchronize.argtypes = [vp]
        L.avbd_get_counts.argtypes = [vp, C.POINTER(AvbdWorld)]
        L.avbd_download.argtypes = [vp, C.POINTER(AvbdWorld)]
        L.avbd_get_status.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32]
        L.avbd_get_trust_region_rho.argtypes = [vp, C.POINTER(C.c_float), C.c_int32]
        L.avbd_get_timings.argtypes = [vp, C.POINTER(AvbdTimings)]
        L.avbd_last_removed.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                                         C.POINTER(C.c_int32)]
        L.avbd_download_bodies.argtypes = [vp, vp, C.c_int32]
        L.avbd_upload_bodies.argtypes = [vp, vp, C.c_int32]
        L.avbd_set_drag.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int32]
        L.avbd_get_colours.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32]
        L.avbd_set_option.argtypes = [vp, C.c_char_p, C.c_int64]
        L.avbd_kernel_stats.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_int32]
        L.avbd_tune_table.argtypes = [vp, C.c_char_p, C.c_int32]
        L.avbd_tune_load.argtypes = [vp, C.c_char_p]
        L.avbd_footprint.argtypes = [C.POINTER(AvbdWorld), C.c_int64, C.POINTER(C.c_int64)]
        L.avbd_get_memory.argtypes = [vp, C.POINTER(C.c_int64)]
        L.avbd_server_create.argtypes = [C.c_int32, C.c_int64, C.c_int32, C.POINTER(vp)]
        L.avbd_server_destroy.argtypes = [vp]
        L.avbd_server_add_scenes.argtypes = [vp, C.POINTER(AvbdWorld), C.POINTER(C.c_int32)]
        L.avbd_server_preload.argtypes = [vp]
        L.avbd_server_now.argtypes = [vp]
        L.avbd_server_now.restype = C.c_double
        L.avbd_server_submit.argtypes = [vp, C.c_int32, C.c_int32, C.c_double, C.POINTER(C.c_int64)]
        L.avbd_server_tick.argtypes = [vp, C.POINTER(C.c_int32)]
        L.avbd_server_request.argtypes = [vp, C.c_int64, C.POINTER(AvbdRequestInfo)]
        L.avbd_server_poll.argtypes = [vp, C.POINTER(AvbdRequestInfo), C.c_int32, C.POINTER(C.c_int32)]
        L.avbd_server_get_stats.argtypes = [vp, C.POINTER(AvbdServerStats)]
        L.avbd_server_download_scene.argtypes = [vp, C.c_int32, C.POINTER(AvbdWorld)]
        L.avbd_server_context.argtypes = [vp]
        L.avbd_server_context.restype = vp
        L.avbd_remove_bodies.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32]
        L.avbd_add_bodies.argtypes = [vp, vp, C.c_int32]
        L.avbd_set_body_colours.argtypes = [vp, vp, C.c_int32]
        L.avbd_get_instances.argtypes = [vp, vp, C.c_int32, C.POINTER(C.c_int32)]
        L.avbd_get_instances_noalloc.argtypes = [vp, vp, C.c_int32, C.POINTER(C.c_int32)]
        L.avbd_get_debug_lines.argtypes = [vp, vp, C.c_int64, C.POINTER(C.c_int64)]
        L.avbd_get_debug_points.argtypes = [vp, vp, C.c_int64, C.POINTER(C.c_int64)]
        L.avbd_snapshot_write.argtypes = [C.c_char_p, C.POINTER(AvbdWorld)]
        L.avbd_snapshot_counts.argtypes = [C.c_char_p, C.POINTER(AvbdWorld)]
        L.avbd_snapshot_read.argtypes = [C.c_char_p, C.POINTER(AvbdWorld)]
        L.avbd_snapshot_save.argtypes = [vp, C.c_char_p]
        L.avbd_snapshot_load.argtypes = [vp, C.c_char_p, C.c_int64]
        _lib = L
    return _lib


def exported_symbols():
    """Function names declared in include/avbd_capi.h."""
    import re

    text = open(HEADER_PATH).read()
    pat = r"^\s*(?:int|double|const char\*|avbd_ctx\*)\s+(avbd_\w+)\s*\("
    return sorted(set(re.findall(pat, text, re.M)))


def footprint(world: World, manifold_capacity: int = 0) -> int:
    """Device bytes an upload of `world` allocates (host-only computation)."""
    keep, cw = world_to_c(world)
    out = C.c_int64()
    _check(lib().avbd_footprint(C.byref(cw), int(manifold_capacity), C.byref(out)))
    return int(out.value)


def save_snapshot(world: World, path) -> None:
    """Write `world` as a snapshot file (avbd_snapshot_write; host only)."""
    keep, cw = world_to_c(world)
    _check(lib().avbd_snapshot_write(os.fsencode(path), C.byref(cw)))


def load_snapshot(path) -> World:
    """Read a snapshot file into a World (avbd_snapshot_read; host only)."""
    cnt = AvbdWorld()
    _check(lib().avbd_snapshot_counts(os.fsencode(path), C.byref(cnt)))
    w, cw, keep = alloc_world(cnt)
    _check(lib().avbd_snapshot_read(os.fsencode(path), C.byref(cw)))
    return w


def _check(rc):
    if rc != 0:
        raise AvbdError(rc, lib().avbd_last_error().decode(errors="replace"))


class Context:
    """A device-resident world on one GPU."""

    def __init__(self, device: int = 0, lanes: int | None = None, fem_lanes: int | None = None, _handle=None):
        self.L = lib()
        self.h = C.c_void_p()
        self._owned = _handle is None
        if _handle is None:
            _check(self.L.avbd_create(device, C.byref(self.h)))
        else:
            self.h = C.c_void_p(_handle)
        if lanes is not None:
            self.set_option("lanes", lanes)
        if fem_lanes is not None:
            self.set_option("fem_lanes", fem_lanes)

    def close(self):
        if self.h and self._owned:
            self.L.avbd_destroy(self.h)
        self.h = C.c_void_p()

    def memory(self) -> int:
        """Device bytes the context holds."""
        out = C.c_int64()
        _check(self.L.avbd_get_memory(self.h, C.byref(out)))
        return int(out.value)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_option(self, name: str, value: int):
        _check(self.L.avbd_set_option(self.h, name.encode(), int(value)))

    def upload(self, world: World, manifold_capacity: int = 0):
        keep, cw = world_to_c(world)
        _check(self.L.avbd_upload(self.h, C.byref(cw), int(manifold_capacity)))

    def save_snapshot(self, path):
        """The device world to a snapshot file (avbd_snapshot_save)."""
        _check(self.L.avbd_snapshot_save(self.h, os.fsencode(path)))

    def load_snapshot(self, path, manifold_capacity: int = 0):
        """A snapshot file into this context (avbd_snapshot_load)."""
        _check(self.L.avbd_snapshot_load(self.h, os.fsencode(path), int(manifold_capacity)))

    def step(self, n: int = 1):
        _check(self.L.avbd_step(self.h, int(n)))

    def step_scenes(self, scene_ids, n: int = 1):
        ids = np.ascontiguousarray(scene_ids, np.int32)
        _check(self.L.avbd_step_scenes(self.h, ids.ctypes.data_as(C.POINTER(C.c_int32)), len(ids), int(n)))

    def synchronize(self):
        _check(self.L.avbd_synchronize(self.h))

    def download(self) -> World:
        cnt = AvbdWorld()
        _check(self.L.avbd_get_counts(self.h, C.byref(cnt)))
        w, cw, keep = alloc_world(cnt)
        _check(self.L.avbd_download(self.h, C.byref(cw)))
        return w

    def status(self, n_scenes: int):
        em = np.zeros(n_scenes, np.int32)
        rem = np.zeros(n_scenes, np.int32)
        _check(self.L.avbd_get_status(self.h, em.ctypes.data_as(C.POINTER(C.c_int32)),
                                      rem.ctypes.data_as(C.POINTER(C.c_int32)), n_scenes))
        return em, rem

    def trust_region_rho(self, n_scenes: int):
        """Solver::trustRegionRho of each scene (avbd_get_trust_region_rho)."""
        rho = np.zeros(n_scenes, np.float32)
        _check(self.L.avbd_get_trust_region_rho(self.h, rho.ctypes.data_as(C.POINTER(C.c_float)), n_scenes))
        return rho

    def remove_bodies(self, bodies):
        """Solver::RemoveBody of live bodies (indices before the call), in order."""
        ids = np.ascontiguousarray(bodies, np.int32)
        _check(self.L.avbd_remove_bodies(self.h, ids.ctypes.data_as(C.POINTER(C.c_int32)), len(ids)))

    def add_bodies(self, bodies: np.ndarray):
        """Solver::AddBody of body records (field scene names the scene), into
        slots removals freed; AvbdError AVBD_E_CAPACITY when a scene has none."""
        b = np.ascontiguousarray(bodies, BODY_DTYPE)
        _check(self.L.avbd_add_bodies(self.h, b.ctypes.data, len(b)))

    def set_body_colours(self, rgb):
        """Mesh::color of every live body ([n, 3]), for the element edge lines."""
        a = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        _check(self.L.avbd_set_body_colours(self.h, a.ctypes.data, len(a)))

    def instances(self, noalloc: bool = False):
        """Renderer instance data per live body ([n, 32]: model matrix 16
        column-major, normal matrix 3 columns of 4, colour 4).  noalloc: the
        entry point that allocates no device memory (at most
        NOALLOC_MAX_BODIES body slots)."""
        f = self.L.avbd_get_instances_noalloc if noalloc else self.L.avbd_get_instances
        n = C.c_int32()
        _check(f(self.h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 32), np.float32)
        _check(f(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def debug_output(self):
        """(lines [n, 10], points [n, 8]) of the last step (option debug_output):
        start xyz, end xyz, rgba / position xyz, 0, rgba."""
        n = C.c_int64()
        _check(self.L.avbd_get_debug_lines(self.h, None, 0, C.byref(n)))
        lines = np.zeros((n.value, 10), np.float32)
        _check(self.L.avbd_get_debug_lines(self.h, lines.ctypes.data, n.value, C.byref(n)))
        _check(self.L.avbd_get_debug_points(self.h, None, 0, C.byref(n)))
        points = np.zeros((n.value, 8), np.float32)
        _check(self.L.avbd_get_debug_points(self.h, points.ctypes.data, n.value, C.byref(n)))
        return lines, points

    def last_removed(self):
        """(scene, index at removal) of the bodies the last step call removed, in order."""
        n = C.c_int32()
        _check(self.L.avbd_last_removed(self.h, None, None, 0, C.byref(n)))
        sc = np.zeros(n.value, np.int32)
        ix = np.zeros(n.value, np.int32)
        _check(self.L.avbd_last_removed(self.h, sc.ctypes.data_as(C.POINTER(C.c_int32)),
                                        ix.ctypes.data_as(C.POINTER(C.c_int32)), n.value, C.byref(n)))
        return list(zip(sc.tolist(), ix.tolist()))

    def download_bodies(self) -> np.ndarray:
        cnt = AvbdWorld()
        _check(self.L.avbd_get_counts(self.h, C.byref(cnt)))
        out = np.zeros(cnt.n_bodies, BODY_DTYPE)
        _check(self.L.avbd_download_bodies(self.h, out.ctypes.data, cnt.n_bodies))
        return out

    def upload_bodies(self, bodies: np.ndarray):
        b = np.ascontiguousarray(bodies, BODY_DTYPE)
        _check(self.L.avbd_upload_bodies(self.h, b.ctypes.data, len(b)))

    def set_drag(self, bodies, added_velocity, dragged=True):
        """Mesh::isDragged / addedDragVelocity of these bodies (live indices):
        the next velocity update adds the velocity once."""
        i = np.ascontiguousarray(bodies, np.int32)
        v = np.ascontiguousarray(np.broadcast_to(np.asarray(added_velocity, np.float32), (len(i), 3)))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dragged, np.int32), (len(i),)))
        _check(self.L.avbd_set_drag(self.h, i.ctypes.data_as(C.POINTER(C.c_int32)),
                                    v.ctypes.data_as(C.POINTER(C.c_float)), d.ctypes.data_as(C.POINTER(C.c_int32)),
                                    len(i)))

    def timings(self) -> dict:
        t = AvbdTimings()
        _check(self.L.avbd_get_timings(self.h, C.byref(t)))
        return {name: getattr(t, name) for name, _ in AvbdTimings._fields_}

    def kernel_stats(self, kernel: str, reset: bool = False):
        """(launches, total device ms) of 'primal' or 'dual' launches since the last reset."""
        n = C.c_int64()
        ms = C.c_double()
        _check(self.L.avbd_kernel_stats(self.h, kernel.encode(), C.byref(n), C.byref(ms), 1 if reset else 0))
        return int(n.value), float(ms.value)

    def tune_table(self) -> dict:
        """Autotuned launch shapes: {(kernel class, log2 bodies per launch):
        (lanes, workgroup size)}."""
        import json

        n = self.L.avbd_tune_table(self.h, None, 0)
        buf = C.create_string_buffer(n + 1)
        self.L.avbd_tune_table(self.h, buf, n + 1)
        raw = json.loads(buf.value.decode() or "{}")
        return {(int(k) // 64, int(k) % 64): tuple(v) if isinstance(v, list) else (v, 256) for k, v in raw.items()}

    def tune_load(self, table: dict):
        import json

        raw = {str(cls * 64 + band): list(v) if isinstance(v, (tuple, list)) else v
               for (cls, band), v in table.items()}
        _check(self.L.avbd_tune_load(self.h, json.dumps(raw).encode()))

    def colours(self, n: int) -> np.ndarray:
        out = np.zeros(n, np.int32)
        _check(self.L.avbd_get_colours(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), n))
        return out
