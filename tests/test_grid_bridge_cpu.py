"""The grid <-> DOF-vector bridge the C++ facade (include/uaamg_openvdb_facade.hpp)
relies on, on host-only handles (no GPU):

- uaamg_poisson_grid_to_vector / _vector_to_grid (gridToVector / vectorToGrid,
  simd_vdb_poisson.cpp:1584-1621) against a brute-force gather over the
  oracle's DOF topology (origins, masks, dofbase), which must equal the
  handle's bit for bit;
- uaamg_viscosity_grid_to_vector / _vector_to_grid (to_Eigenvector /
  write_to_FloatGrid3, simd_viscosity3d.cpp:2089-2139) likewise per channel;
- uaamg_viscosity_build_rhs_packed on packed_FloatGrid3's three channels
  equals the oracle's build_rhs of the Vec3f velocity;
- uaamg_packed3_from_vec3 / _to_vec3 (packed3grids.cpp:11-83) against a
  brute-force restatement of the topology-copy and topology-union rules,
  with negative origins, partial active masks and three different channel
  topologies, and the round trip.
"""
import ctypes as C

import numpy as np

import oracle as O
import uaamg_amd as U
from tests._brute import tree_key
from uaamg_amd.viscosity import L_with_level, ViscScene


def _grid(origins, values, active=None, bg=(0.0, 0.0, 0.0)):
    return U.Grid(np.ascontiguousarray(origins, np.int32), np.ascontiguousarray(values, np.float32),
                  None if active is None else np.ascontiguousarray(active, np.uint64), tuple(bg))


def _bits(mask8):
    """[8] uint64 words -> [512] bool, bit b of word w = offset 64w + b"""
    return np.unpackbits(np.ascontiguousarray(mask8, np.uint64).view(np.uint8), bitorder="little").astype(bool)


def _gather(org, mask, base, grid_org, grid_vals):
    """brute force: DOF d of leaf i = i's base + rank of the voxel among i's DOF voxels"""
    where = {tuple(int(v) for v in o): k for k, o in enumerate(grid_org)}
    out = np.full(int(sum(_bits(m).sum() for m in mask)), np.nan, np.float32)
    for i, o in enumerate(org):
        on = np.nonzero(_bits(mask[i]))[0]
        k = where.get(tuple(int(v) for v in o))
        out[base[i]:base[i] + on.size] = 0.0 if k is None else grid_vals[k][on]
    return out


def test_poisson_grid_vector_bridge():
    rng = np.random.default_rng(3)
    L = U.lib()
    for cfg, n, seed in ((0, 32, 0), (1, 48, 1)):
        s = U.Scene(cfg, n, seed=seed)
        lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
        orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
        org, mask, base = orc.topology(0)
        org2, mask2, base2 = lap.level_topology(0)
        assert np.array_equal(org, org2) and np.array_equal(mask, mask2) and np.array_equal(base, base2)
        # a grid on the phi topology with random values; a quarter of its leaves dropped (they read 0)
        keep = rng.random(s.phi.origins.shape[0]) > 0.25
        go = s.phi.origins[keep]
        gv = rng.standard_normal((go.shape[0], 512)).astype(np.float32)
        g = _grid(go, gv)
        vec = np.full(lap.mNumDof, np.nan, np.float32)
        assert L.uaamg_poisson_grid_to_vector(lap._h, C.byref(g._struct()), vec.ctypes.data_as(C.c_void_p)) == 0
        assert np.array_equal(vec, _gather(org, mask, base, go, gv))
        # vector -> grid on the DOF topology, and back
        x = rng.standard_normal(lap.mNumDof).astype(np.float32)
        br = np.full((org.shape[0], 512), np.nan, np.float32)
        assert L.uaamg_poisson_vector_to_grid(lap._h, x.ctypes.data_as(C.c_void_p), br.ctypes.data_as(C.c_void_p)) == 0
        for i in range(org.shape[0]):
            on = _bits(mask[i])
            assert np.array_equal(br[i][on], x[base[i]:base[i] + on.sum()])
            assert not br[i][~on].any()
        back = np.zeros_like(x)
        gb = _grid(org, br, mask)
        assert L.uaamg_poisson_grid_to_vector(lap._h, C.byref(gb._struct()), back.ctypes.data_as(C.c_void_p)) == 0
        assert np.array_equal(back, x)


def test_viscosity_grid_vector_bridge_and_packed_rhs():
    rng = np.random.default_rng(5)
    L = U.lib()
    v = ViscScene(0, 24, seed=0, mu=1.0)
    lw = L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx, device=-1)
    ov = O.OracleViscosity(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx)
    org, mask, base = ov.topology(0)
    org2, mask2, base2 = lw.level_topology(0)
    assert np.array_equal(org, org2) and np.array_equal(mask, mask2) and np.array_equal(base, base2)
    ndof = lw.m_ndof
    # packed_FloatGrid3 of the scene's velocity, then the right-hand side from the three channels
    vel = v.velocity
    nl = vel.origins.shape[0]
    ch = np.zeros((3, nl, 512), np.float32)
    assert L.uaamg_packed3_from_vec3(C.byref(vel._struct()), ch.ctypes.data_as(C.c_void_p)) == 0
    grids = [_grid(vel.origins, ch[c], vel.active, (vel.background[c],) * 3) for c in range(3)]
    rhs = np.zeros(ndof, np.float32)
    st = [g._struct() for g in grids]
    assert L.uaamg_viscosity_build_rhs_packed(lw._h, *[C.byref(x) for x in st], None,
                                              rhs.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(rhs, ov.build_rhs(vel, None))
    # to_Eigenvector of three random channel grids (some leaves missing per channel)
    gs, vals, orgs = [], [], []
    for c in range(3):
        keep = rng.random(nl) > 0.2
        orgs.append(vel.origins[keep])
        vals.append(rng.standard_normal((int(keep.sum()), 512)).astype(np.float32))
        gs.append(_grid(orgs[-1], vals[-1]))
    vec = np.full(ndof, np.nan, np.float32)
    st = [g._struct() for g in gs]
    assert L.uaamg_viscosity_grid_to_vector(lw._h, *[C.byref(x) for x in st], vec.ctypes.data_as(C.c_void_p)) == 0
    ref = np.full(ndof, np.nan, np.float32)
    for c in range(3):
        where = {tuple(int(t) for t in o): k for k, o in enumerate(orgs[c])}
        for i, o in enumerate(org):
            if base[i, c] < 0:
                continue
            on = np.nonzero(_bits(mask[i, c]))[0]
            k = where.get(tuple(int(t) for t in o))
            ref[base[i, c]:base[i, c] + on.size] = 0.0 if k is None else vals[c][k][on]
    assert np.array_equal(vec, ref)
    # write_to_FloatGrid3 onto the DOF topology and back
    x = rng.standard_normal(ndof).astype(np.float32)
    br = np.full((org.shape[0], 3, 512), np.nan, np.float32)
    assert L.uaamg_viscosity_vector_to_grid(lw._h, x.ctypes.data_as(C.c_void_p), br.ctypes.data_as(C.c_void_p)) == 0
    for i in range(org.shape[0]):
        for c in range(3):
            on = _bits(mask[i, c])
            if base[i, c] >= 0:
                assert np.array_equal(br[i, c][on], x[base[i, c]:base[i, c] + on.sum()])
            assert not br[i, c][~on].any()
    back = np.zeros_like(x)
    gb = [_grid(org, br[:, c], mask[:, c]) for c in range(3)]   # kept alive while the structs are used
    st = [g._struct() for g in gb]
    assert L.uaamg_viscosity_grid_to_vector(lw._h, *[C.byref(t) for t in st], back.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(back, x)


def _random_vec3(rng, nleaf, bg):
    cells = rng.choice(np.arange(-3, 3), size=(nleaf * 3, 3))
    org = np.unique(cells * 8, axis=0)[:nleaf]
    rng.shuffle(org)
    vals = rng.standard_normal((org.shape[0], 512, 3)).astype(np.float32)
    act = rng.integers(0, 2**63, size=(org.shape[0], 8), dtype=np.uint64)
    return _grid(org, vals, act, bg)


def test_packed3_from_vec3():
    rng = np.random.default_rng(11)
    g = _random_vec3(rng, 9, (0.5, -1.0, 2.0))
    out = np.full((3, g.origins.shape[0], 512), np.nan, np.float32)
    assert U.lib().uaamg_packed3_from_vec3(C.byref(g._struct()), out.ctypes.data_as(C.c_void_p)) == 0
    for i in range(g.origins.shape[0]):
        on = _bits(g.active[i])
        for c in range(3):
            exp = np.where(on, g.values[i, :, c], np.float32(g.background[c]))
            assert np.array_equal(out[c, i], exp)


def test_packed3_to_vec3_and_round_trip():
    rng = np.random.default_rng(12)
    L = U.lib()
    chans = []
    for c in range(3):
        g = _random_vec3(rng, 5 + c, (0.0, 0.0, 0.0))
        chans.append(_grid(g.origins, g.values[:, :, 0], g.active, (float(c),) * 3))
    st = [g._struct() for g in chans]
    n = C.c_int64(0)
    assert L.uaamg_packed3_to_vec3(*[C.byref(t) for t in st], C.byref(n), None, None, None) == 0
    org = np.zeros((n.value, 3), np.int32)
    act = np.zeros((n.value, 8), np.uint64)
    val = np.full((n.value, 512, 3), np.nan, np.float32)
    assert L.uaamg_packed3_to_vec3(*[C.byref(t) for t in st], C.byref(n), org.ctypes.data_as(C.c_void_p),
                                   act.ctypes.data_as(C.c_void_p), val.ctypes.data_as(C.c_void_p)) == 0
    # brute force: union of the leaves in tree order, OR of the masks, channel values where active
    union = sorted({tuple(int(t) for t in o) for g in chans for o in g.origins}, key=tree_key)
    assert [tuple(o) for o in org.tolist()] == union
    for i, o in enumerate(union):
        m = np.zeros(512, bool)
        exp = np.zeros((512, 3), np.float32)
        for c, g in enumerate(chans):
            hit = [k for k, q in enumerate(g.origins) if tuple(int(t) for t in q) == o]
            if not hit:
                continue
            on = _bits(g.active[hit[0]])
            m |= on
            exp[on, c] = g.values[hit[0]][on]
        assert np.array_equal(_bits(act[i]), m)
        assert np.array_equal(val[i], exp)
    # round trip: a Vec3 grid in tree order with zero background comes back on its active voxels
    g = _random_vec3(rng, 7, (0.0, 0.0, 0.0))
    order = sorted(range(g.origins.shape[0]), key=lambda k: tree_key(g.origins[k]))
    g = _grid(g.origins[order], g.values[order], g.active[order])
    ch = np.zeros((3, g.origins.shape[0], 512), np.float32)
    assert L.uaamg_packed3_from_vec3(C.byref(g._struct()), ch.ctypes.data_as(C.c_void_p)) == 0
    gb = [_grid(g.origins, ch[c], g.active) for c in range(3)]
    st = [x._struct() for x in gb]
    n = C.c_int64(g.origins.shape[0])
    org = np.zeros((n.value, 3), np.int32)
    act = np.zeros((n.value, 8), np.uint64)
    val = np.zeros((n.value, 512, 3), np.float32)
    assert L.uaamg_packed3_to_vec3(*[C.byref(t) for t in st], C.byref(n), org.ctypes.data_as(C.c_void_p),
                                   act.ctypes.data_as(C.c_void_p), val.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(org, g.origins) and np.array_equal(act, g.active)
    for i in range(n.value):
        on = _bits(act[i])
        assert np.array_equal(val[i][on], g.values[i][on])
        assert not val[i][~on].any()
