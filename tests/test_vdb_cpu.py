"""The OpenVDB file reader (uaamg_vdb_*), on the CPU.

The reference's scene file tests/golden/sphere_source.vdb (Houdini 19, Blosc
with the LZ4 codec, byte shuffle and active-mask compression) pins the whole
decode path against what the file itself states and what its geometry fixes:
- the active voxel count and active bounding box equal the file's own
  metadata (file_voxel_count, file_bbox_min / file_bbox_max, written by the
  exporter from its in-memory grid);
- the active values are a signed distance to one sphere (least-squares fit:
  RMS deviation below 0.05 voxel), the inactive values are exactly +-background
  (3 voxels, the narrow band), and the interior tiles lie inside the sphere;
- the leaves come in OpenVDB tree order.
The Blosc and LZ4 decoders are also checked on buffers made here from their
published formats (a greedy LZ4 encoder, the Blosc 1 container with and
without shuffle, splits and short last blocks, the zlib codec, buffers
stored uncompressed), and malformed inputs fail cleanly.
"""
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

import uaamg_amd as U
from uaamg_amd.vdb import blosc_decompress, read_vdb

FIXTURE = Path(__file__).resolve().parent / "golden" / "sphere_source.vdb"


@pytest.fixture(scope="module")
def source():
    grids = read_vdb(str(FIXTURE))
    assert len(grids) == 1
    return grids[0]


def _active_points(g):
    G = g.grid
    n = G.origins.shape[0]
    bits = np.unpackbits(G.active.view(np.uint8).reshape(n, 64), axis=1, bitorder="little").astype(bool)
    off = np.arange(512)
    local = np.stack([off >> 6, (off >> 3) & 7, off & 7], -1)
    ijk = G.origins[:, None, :] + local[None]
    return bits, ijk


def test_grid_header_and_metadata(source):
    g = source
    assert (g.name, g.type, g.grid_class, g.ncomp, g.half_float) == ("source", "Tree_float_5_4_3", "level set", 1,
                                                                     False)
    assert g.voxel_size == pytest.approx((0.00025,) * 3, rel=1e-6)
    assert g.translation == (0.0, 0.0, 0.0)
    # a level set's background is its narrow-band half width: 3 voxels
    assert g.background[0] == pytest.approx(3 * g.voxel_size[0], rel=1e-6)
    bits, ijk = _active_points(g)
    assert int(bits.sum()) == g.meta_voxel_count == 30346
    act = ijk[bits]
    assert tuple(act.min(0)) + tuple(act.max(0)) == g.meta_bbox
    assert np.all(g.grid.origins % 8 == 0)


def test_values_are_a_sphere_distance(source):
    from scipy.optimize import least_squares
    g = source
    h = g.voxel_size[0]
    bits, ijk = _active_points(g)
    pts = ijk[bits] * h
    phi = g.grid.values[bits]
    fit = least_squares(lambda p: np.linalg.norm(pts - p[:3], axis=1) - p[3] - phi, np.r_[pts.mean(0), 10 * h])
    centre, radius = fit.x[:3] / h, fit.x[3] / h
    rms = np.sqrt(np.mean(fit.fun ** 2)) / h
    assert np.allclose(centre, (0, 1000, 0), atol=0.01)
    assert abs(radius - 20.0) < 0.1
    assert rms < 0.05
    assert np.abs(phi).max() <= g.background[0]
    # inactive voxels hold exactly -background inside and +background outside
    inact = g.grid.values[~bits]
    assert set(np.unique(inact).tolist()) == {-g.background[0], g.background[0]}
    d = np.linalg.norm(ijk[~bits] * h - fit.x[:3], axis=1) - fit.x[3]
    assert np.all(np.sign(inact) == np.sign(d))
    # interior tiles: 8^3 blocks of -background, inactive, inside the sphere
    to, td, tv, ta = g.tiles
    assert len(td) > 0 and np.all(td == 8) and not ta.any()
    assert np.all(tv[:, 0] == -g.background[0])
    assert np.all(np.linalg.norm((to + 4) * h - fit.x[:3], axis=1) < fit.x[3] - 4 * h)


def test_leaves_in_tree_order(source):
    o = source.grid.origins.astype(np.int64)
    root = o >> 12
    i5 = (o >> 7) & 31
    i4 = (o >> 3) & 15
    key = [tuple(r) + (int(a[0] << 10 | a[1] << 5 | a[2]), int(b[0] << 8 | b[1] << 4 | b[2]))
           for r, a, b in zip(root, i5, i4)]
    assert key == sorted(key) and len(set(key)) == len(key)


def test_expanded_tiles(source):
    """expand_tiles=128 turns the 8 interior tiles into inactive leaves of -background, keeps every
    original leaf as it was and the tree order."""
    ex = read_vdb(str(FIXTURE), expand_tiles=128)[0]
    a, b = source.grid, ex.grid
    assert b.origins.shape[0] == a.origins.shape[0] + len(source.tiles[1])
    assert len(ex.tiles[1]) == 0
    where = {tuple(o): i for i, o in enumerate(a.origins)}
    new = []
    for j, o in enumerate(b.origins):
        i = where.get(tuple(o))
        if i is None:
            new.append(j)
        else:
            assert np.array_equal(a.values[i], b.values[j]) and np.array_equal(a.active[i], b.active[j])
    assert sorted(tuple(o) for o in b.origins[new]) == sorted(tuple(o) for o in source.tiles[0])
    assert np.all(b.values[new] == -source.background[0]) and not b.active[new].any()
    o = b.origins.astype(np.int64)
    key = [tuple(r) + (int(p[0] << 10 | p[1] << 5 | p[2]), int(q[0] << 8 | q[1] << 4 | q[2]))
           for r, p, q in zip(o >> 12, (o >> 7) & 31, (o >> 3) & 15)]
    assert key == sorted(key)
    with pytest.raises(U.UaamgError):   # the leaf budget is checked before anything changes
        read_vdb(str(FIXTURE), expand_tiles=128, max_leaves=160)


# ---- Blosc / LZ4 on buffers made from the published formats

def _lz4_compress(data: bytes) -> bytes:
    """Greedy LZ4 block encoder (4-byte hash of the last position; the last 5
    bytes stay literals, a match starts at least 12 bytes before the end)."""
    n = len(data)
    out = bytearray()
    table = {}
    i = anchor = 0

    def emit(lit: bytes, offset: int = 0, ml: int = 0):
        ln = len(lit)
        tok_l = min(ln, 15)
        tok_m = min(ml - 4, 15) if ml else 0
        out.append((tok_l << 4) | tok_m)
        if ln >= 15:
            r = ln - 15
            while r >= 255:
                out.append(255)
                r -= 255
            out.append(r)
        out.extend(lit)
        if ml:
            out.extend(struct.pack("<H", offset))
            if ml - 4 >= 15:
                r = ml - 4 - 15
                while r >= 255:
                    out.append(255)
                    r -= 255
                out.append(r)

    while i < n - 12:
        key = data[i:i + 4]
        j = table.get(key)
        table[key] = i
        if j is not None and i - j <= 65535:
            ml = 4
            while i + ml < n - 5 and data[j + ml] == data[i + ml]:
                ml += 1
            emit(data[anchor:i], i - j, ml)
            i += ml
            anchor = i
        else:
            i += 1
    emit(data[anchor:])
    return bytes(out)


def _blosc_pack(data: bytes, typesize: int, blocksize: int, codec: str = "lz4", shuffle: bool = True,
                no_split: bool = False, memcpy: bool = False) -> bytes:
    nbytes = len(data)
    cid = {"lz4": 1, "zlib": 3}[codec]
    flags = (1 if shuffle else 0) | (cid << 5) | (0x10 if no_split else 0) | (0x2 if memcpy else 0)
    if memcpy:
        return struct.pack("<BBBBIII", 2, 1, flags, typesize, nbytes, blocksize, 16 + nbytes) + data
    nblocks = (nbytes + blocksize - 1) // blocksize
    leftover = nbytes % blocksize
    body = bytearray()
    starts = []
    base = 16 + 4 * nblocks
    for j in range(nblocks):
        last_short = j == nblocks - 1 and leftover
        bsize = leftover if last_short else blocksize
        blk = data[j * blocksize:j * blocksize + bsize]
        if shuffle and typesize > 1:
            ne = bsize // typesize
            a = np.frombuffer(blk[:ne * typesize], np.uint8).reshape(ne, typesize)
            blk = a.T.tobytes() + blk[ne * typesize:]
        split = (not no_split) and typesize <= 16 and bsize // typesize >= 128 and bsize % typesize == 0 \
            and not last_short
        ns = typesize if split else 1
        neb = bsize // ns
        starts.append(base + len(body))
        for k in range(ns):
            chunk = blk[k * neb:(k + 1) * neb]
            c = _lz4_compress(chunk) if codec == "lz4" else zlib.compress(chunk)
            if len(c) >= neb:
                c = chunk
            body += struct.pack("<I", len(c)) + c
    table = b"".join(struct.pack("<I", s) for s in starts)
    return struct.pack("<BBBBIII", 2, 1, flags, typesize, nbytes, blocksize, base + len(body)) + table + bytes(body)


def _payloads():
    rng = np.random.default_rng(5)
    smooth = np.cumsum(rng.standard_normal(3000).astype(np.float32) * 1e-3).astype(np.float32).tobytes()
    rep = (b"abcdefgh" * 700)[:5555]
    noise = rng.integers(0, 256, 4100, dtype=np.uint8).tobytes()
    runs = bytes([7]) * 3000 + bytes(range(256)) * 3
    return {"smooth": smooth, "repetitive": rep, "noise": noise, "runs": runs, "small": smooth[:68]}


@pytest.mark.parametrize("name", ["smooth", "repetitive", "noise", "runs", "small"])
@pytest.mark.parametrize("codec,shuffle,no_split,blocksize",
                         [("lz4", True, False, 512), ("lz4", True, False, 4096), ("lz4", False, False, 1024),
                          ("lz4", True, True, 2048), ("zlib", True, False, 512), ("zlib", False, True, 3000)])
def test_blosc_round_trip(name, codec, shuffle, no_split, blocksize):
    data = _payloads()[name]
    buf = _blosc_pack(data, 4, blocksize, codec, shuffle, no_split)
    assert blosc_decompress(buf, len(data)) == data


def test_blosc_stored_uncompressed():
    data = _payloads()["noise"]
    assert blosc_decompress(_blosc_pack(data, 4, 512, memcpy=True), len(data)) == data


def test_lz4_overlapping_match():
    # literal "ab", then a 20-byte match at offset 2 (copies its own output), then literals "xyz!!"
    stream = bytes([(2 << 4) | 15]) + b"ab" + struct.pack("<H", 2) + bytes([1]) + bytes([5 << 4]) + b"xyz!!"
    data = b"ab" * 11 + b"xyz!!"
    hdr = struct.pack("<BBBBIII", 2, 1, (1 << 5) | 0x10, 1, len(data), len(data), 0)
    body = struct.pack("<I", 16 + 4) + struct.pack("<I", len(stream)) + stream
    buf = bytearray(hdr + body)
    buf[12:16] = struct.pack("<I", len(buf))
    assert blosc_decompress(bytes(buf), len(data)) == data


def test_malformed_buffers_fail_cleanly(tmp_path):
    data = _payloads()["smooth"]
    good = _blosc_pack(data, 4, 512)
    for bad in (good[:10], good[:-7], good[:16] + b"\xff" * (len(good) - 16)):
        with pytest.raises(U.UaamgError):
            blosc_decompress(bad, len(data))
    with pytest.raises(U.UaamgError):      # wrong expected size
        blosc_decompress(good, len(data) + 4)
    raw = FIXTURE.read_bytes()
    for cut in (10, 200, len(raw) // 2, len(raw) - 1):
        p = tmp_path / f"cut{cut}.vdb"
        p.write_bytes(raw[:cut])
        with pytest.raises(U.UaamgError):
            read_vdb(str(p))
    p = tmp_path / "notvdb.vdb"
    p.write_bytes(b"\0" * 64)
    with pytest.raises(U.UaamgError):
        read_vdb(str(p))
