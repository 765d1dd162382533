"""OpenVDB files without OpenVDB (uaamg_vdb_*): the grids FLIP reads with
openvdb::io::File (its liquid source / collision level sets and velocity
grids, FLIP_vdb/main.cpp:108-131,161-185) and writes each frame
(sdfvel<n>.vdb, FLIP_vdb/main.cpp:266-284), as flattened leaf grids the
solvers take."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import Grid, _check, _Grid, lib


class _GridDesc(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("type", C.c_char_p),
        ("grid_class", C.c_char_p),
        ("decoded", C.c_int32),
        ("ncomp", C.c_int32),
        ("half_float", C.c_int32),
        ("nleaf", C.c_int64),
        ("ntiles", C.c_int64),
        ("voxel_size", C.c_double * 3),
        ("translation", C.c_double * 3),
        ("background", C.c_float * 3),
        ("meta_voxel_count", C.c_int64),
        ("meta_bbox", C.c_int32 * 6),
        ("has_meta_bbox", C.c_int32),
    ]


_bound = False


def _bind():
    global _bound
    L = lib()
    if not _bound:
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        sig = {
            "uaamg_vdb_open": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
            "uaamg_vdb_close": (None, [vp]),
            "uaamg_vdb_grid_count": (i32, [vp]),
            "uaamg_vdb_grid_info": (C.c_int, [vp, i32, C.POINTER(_GridDesc)]),
            "uaamg_vdb_grid": (C.c_int, [vp, i32, C.POINTER(_Grid)]),
            "uaamg_vdb_grid_tiles": (C.c_int, [vp, i32, vp, vp, vp, vp]),
            "uaamg_blosc_decompress": (C.c_int, [vp, i64, vp, i64]),
            "uaamg_vdb_expand_tiles": (C.c_int, [vp, i32, i32, i64, C.POINTER(i64)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


@dataclass
class VdbGrid:
    """One grid of the file: `grid` holds its leaves (None for grid types that are not decoded),
    `tiles` the root / internal tiles that are active or differ from the background as
    (origins [n,3], edge [n], values [n,3], active [n])."""
    name: str
    type: str
    grid_class: str
    ncomp: int
    half_float: bool
    voxel_size: tuple
    translation: tuple
    background: tuple
    meta_voxel_count: int
    meta_bbox: tuple | None = None
    grid: Grid | None = None
    tiles: tuple = field(default_factory=tuple)


def read_vdb(path: str, expand_tiles: int = 0, max_leaves: int = 1 << 24) -> list[VdbGrid]:
    """Every grid of an OpenVDB file, in file order. expand_tiles = 8, 128 or 4096 turns the tiles of
    at most that edge into leaves holding the tile value (what an OpenVDB accessor returns there)."""
    L = _bind()
    h = C.c_void_p()
    _check(L.uaamg_vdb_open(str(path).encode(), C.byref(h)), "uaamg_vdb_open")
    try:
        out = []
        for i in range(L.uaamg_vdb_grid_count(h)):
            if expand_tiles:
                d = _GridDesc()
                _check(L.uaamg_vdb_grid_info(h, i, C.byref(d)), "uaamg_vdb_grid_info")
                if d.decoded:
                    added = C.c_int64()
                    _check(L.uaamg_vdb_expand_tiles(h, i, int(expand_tiles), int(max_leaves), C.byref(added)),
                           "uaamg_vdb_expand_tiles")
            d = _GridDesc()
            _check(L.uaamg_vdb_grid_info(h, i, C.byref(d)), "uaamg_vdb_grid_info")
            g = VdbGrid(d.name.decode(), d.type.decode(), d.grid_class.decode(), d.ncomp, bool(d.half_float),
                        tuple(d.voxel_size), tuple(d.translation), tuple(d.background[:max(d.ncomp, 1)]),
                        int(d.meta_voxel_count), tuple(d.meta_bbox) if d.has_meta_bbox else None)
            if d.decoded:
                s = _Grid()
                _check(L.uaamg_vdb_grid(h, i, C.byref(s)), "uaamg_vdb_grid")
                n = int(s.nleaf)
                org = np.ctypeslib.as_array(C.cast(s.origins, C.POINTER(C.c_int32)), shape=(n * 3,)).reshape(n, 3)
                act = np.ctypeslib.as_array(C.cast(s.active, C.POINTER(C.c_uint64)), shape=(n * 8,)).reshape(n, 8)
                nv = n * 512 * d.ncomp
                val = np.ctypeslib.as_array(C.cast(s.values, C.POINTER(C.c_float)), shape=(nv,))
                shape = (n, 512) if d.ncomp == 1 else (n, 512, 3)
                bg = tuple(d.background[:d.ncomp]) if d.ncomp == 3 else float(d.background[0])
                g.grid = Grid(org.copy(), val.reshape(shape).copy(), act.copy(), bg)
                nt = int(d.ntiles)
                to = np.zeros((nt, 3), np.int32)
                td = np.zeros(nt, np.int32)
                tv = np.zeros((nt, 3), np.float32)
                ta = np.zeros(nt, np.int32)
                if nt:
                    _check(L.uaamg_vdb_grid_tiles(h, i, to.ctypes.data, td.ctypes.data, tv.ctypes.data, ta.ctypes.data),
                           "uaamg_vdb_grid_tiles")
                g.tiles = (to, td, tv, ta.astype(bool))
            out.append(g)
        return out
    finally:
        L.uaamg_vdb_close(h)


def blosc_decompress(buf: bytes, nbytes: int) -> bytes:
    """The reader's Blosc 1 decoder (LZ4 or zlib codec, byte shuffle) on one buffer."""
    L = _bind()
    src = np.frombuffer(bytes(buf), np.uint8)
    out = np.zeros(nbytes, np.uint8)
    _check(L.uaamg_blosc_decompress(src.ctypes.data, src.size, out.ctypes.data, nbytes), "uaamg_blosc_decompress")
    return out.tobytes()
