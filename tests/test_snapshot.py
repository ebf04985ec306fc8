"""World snapshot files (avbd_snapshot_*): the ABI records of a world in one
binary file.  Host-only: round trips are byte-exact for every record kind,
the header says what it should, and missing, foreign, truncated and corrupt
files are refused with AVBD_E_IO.  (The device save/load round trip is in
test_gpu_snapshot.py.)
"""
import os

import numpy as np
import pytest

import oracle_py as O
from skyephyx_amd import abi, capi, scenes
from skyephyx_amd.world import World

FIELDS = ("scenes", "bodies", "joints", "springs", "manifolds", "tets", "tris")
AVBD_E_IO = -7


def _warm(world, steps):
    o = O.Oracle(world, fp64=False, order=O.NATURAL)
    o.step(steps)
    return o.save()


WORLDS = {
    "pyramid_settled": lambda: _warm(scenes.pyramid_2d(), 20),  # manifolds with contacts
    "spring_bridge": lambda: _warm(scenes.spring_bridge(4), 5),
    "fem_block": lambda: scenes.fem_block(n=3, cell=0.5),
    "cloth": lambda: scenes.cloth_sheet(n=3, spacing=0.5),
    "lattice_mix": lambda: _warm(scenes.box_lattice(shape=(3, 2, 2), spacing=3.0, n_chain=4, n_springs=3,
                                                    bounds=50.0), 10),
    "several_scenes": lambda: World.concat([scenes.pyramid_2d(rows=3), scenes.joint_chain(3),
                                            scenes.fem_block(n=2, cell=0.5)]),
}


def _same(a: World, b: World):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert x.dtype == y.dtype and len(x) == len(y), f
        assert x.tobytes() == y.tobytes(), f


@pytest.mark.parametrize("name", sorted(WORLDS))
def test_round_trip_is_byte_exact(name, tmp_path):
    w = WORLDS[name]()
    path = tmp_path / f"{name}.avbdsnap"
    capi.save_snapshot(w, path)
    _same(capi.load_snapshot(path), w)


def test_header_and_layout(tmp_path):
    w = _warm(scenes.pyramid_2d(), 5)
    path = tmp_path / "p.avbdsnap"
    capi.save_snapshot(w, path)
    raw = np.fromfile(path, np.uint8)
    assert raw[:8].tobytes() == b"AVBDSNAP"
    u32 = raw[8:56].view(np.uint32)
    assert u32[0] == 1 and u32[1] == 128 and u32[2] == capi.lib().avbd_abi_version() and u32[3] == 7
    sizes = [getattr(w, f).dtype.itemsize for f in FIELDS]
    assert list(u32[4:11]) == sizes
    counts = raw[56:112].view(np.int64)
    assert list(counts) == [len(getattr(w, f)) for f in FIELDS]
    # sections at 64-byte aligned offsets, records back to back
    at = end = 128
    for f, size in zip(FIELDS, sizes):
        at = (at + 63) // 64 * 64
        n = len(getattr(w, f))
        assert raw[at:at + n * size].tobytes() == getattr(w, f).tobytes(), f
        at += n * size
        if n:
            end = at
    assert len(raw) == end  # the file ends with the last record


def _expect_io_error(path):
    with pytest.raises(capi.AvbdError) as e:
        capi.load_snapshot(path)
    assert e.value.code == AVBD_E_IO


def test_bad_files_are_refused(tmp_path):
    w = scenes.pyramid_2d(rows=3)
    good = tmp_path / "good.avbdsnap"
    capi.save_snapshot(w, good)
    raw = bytearray(open(good, "rb").read())

    _expect_io_error(tmp_path / "missing.avbdsnap")

    bad = tmp_path / "magic.avbdsnap"
    b = bytearray(raw)
    b[0:8] = b"NOTASNAP"
    bad.write_bytes(bytes(b))
    _expect_io_error(bad)

    bad = tmp_path / "version.avbdsnap"
    b = bytearray(raw)
    b[8] = 9
    bad.write_bytes(bytes(b))
    _expect_io_error(bad)

    bad = tmp_path / "layout.avbdsnap"  # a scene record of another size
    b = bytearray(raw)
    b[24] ^= 4
    bad.write_bytes(bytes(b))
    _expect_io_error(bad)

    bad = tmp_path / "short.avbdsnap"
    bad.write_bytes(bytes(raw[:-10]))
    _expect_io_error(bad)

    bad = tmp_path / "trailing.avbdsnap"
    bad.write_bytes(bytes(raw) + b"\0" * 64)
    _expect_io_error(bad)

    bad = tmp_path / "header_only.avbdsnap"
    bad.write_bytes(bytes(raw[:100]))
    _expect_io_error(bad)

    bad = tmp_path / "corrupt.avbdsnap"  # one flipped bit in a body position
    b = bytearray(raw)
    b[len(b) - 50] ^= 1
    bad.write_bytes(bytes(b))
    _expect_io_error(bad)


def test_empty_sections(tmp_path):
    w = scenes.mass_springs(2)  # no manifolds, tets or tris
    path = tmp_path / "ms.avbdsnap"
    capi.save_snapshot(w, path)
    _same(capi.load_snapshot(path), w)
    cnt = abi.AvbdWorld()
    capi._check(capi.lib().avbd_snapshot_counts(os.fsencode(path), cnt))
    assert (cnt.n_manifolds, cnt.n_tets, cnt.n_tris) == (0, 0, 0)
    assert cnt.n_bodies == len(w.bodies)
