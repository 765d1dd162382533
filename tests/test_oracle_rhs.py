"""Pins the oracle's right-hand side (BuildPoissonRhs, simd_vdb_poisson.cpp:19-93) on the CPU.

b(g) = sum over channels c of -(1/dx) * (F_c(g + e_c) - F_c(g)), F_c = w_c u_c + (1 - w_c) us_c, the
face value of channel c at voxel g being the face on its minus side; b = 0 where all six face
weights are 0. The checks are closed forms that a plausible slip breaks:
  - uniform fields have zero divergence exactly;
  - a linear liquid velocity gives minus the sum of its slopes (a wrong sign flips it);
  - a quadratic field gives -(2 g + 1) (reading the + face at g instead of g + e_c gives -(2 g - 1));
  - a fractional weight blends two linear fields, -(w a + (1 - w) b) per axis (a dropped solid term
    or a swapped w / 1 - w fails);
  - all-zero weights give 0 whatever the solid velocity (the hasNonZeroWeight rule);
and a dense brute-force divergence over random fields on random topologies with negative origins
(missing leaves read the grids' backgrounds).
"""
from dataclasses import dataclass

import numpy as np
import pytest

import oracle as O
from tests import _brute as B


@dataclass
class G:
    origins: np.ndarray
    values: np.ndarray
    active: np.ndarray | None
    background: tuple


DX = 0.125


def _leaf_box(lo, hi):
    r = range(lo, hi, 8)
    return np.array([(x, y, z) for x in r for y in r for z in r], np.int32)


def _voxel_coords(org):
    off = np.arange(512)
    lx, ly, lz = off >> 6, (off >> 3) & 7, off & 7
    return (org[:, 0:1] + lx[None, :], org[:, 1:2] + ly[None, :], org[:, 2:3] + lz[None, :])


def _field(org, f):
    gx, gy, gz = _voxel_coords(org)
    return np.ascontiguousarray(f(gx, gy, gz), np.float32)


def _vec_grid(org, fx, fy, fz, bg):
    gx, gy, gz = _voxel_coords(org)
    v = np.stack([fx(gx, gy, gz), fy(gx, gy, gz), fz(gx, gy, gz)], axis=-1)
    return G(org, np.ascontiguousarray(np.broadcast_to(v, (org.shape[0], 512, 3)), np.float32), None, bg)


def _liquid_block(lo=-8, hi=8, w=1.0):
    """All-liquid leaves over [lo, hi)^3 with uniform face weight w: every voxel is a DOF."""
    org = _leaf_box(lo, hi)
    phi = G(org, -np.ones((org.shape[0], 512), np.float32), None, (-1.0,))
    wg = _vec_grid(org, *(lambda x, y, z: np.full(x.shape, w),) * 3, (w, w, w))
    return org, phi, wg


def _dof_coords(orc):
    org, mask, base = orc.topology(0)
    coords = np.zeros((orc.ndof(0), 3), np.int64)
    for i in range(org.shape[0]):
        bits = np.unpackbits(mask[i].view(np.uint8), bitorder="little").astype(bool)
        offs = np.nonzero(bits)[0]
        coords[base[i]:base[i] + offs.size] = np.stack([org[i, 0] + (offs >> 6), org[i, 1] + ((offs >> 3) & 7),
                                                        org[i, 2] + (offs & 7)], axis=1)
    return coords


def _vel(org_v, fx, fy, fz, bg=0.0):
    return (G(org_v, _field(org_v, fx), None, (bg,)), G(org_v, _field(org_v, fy), None, (bg,)),
            G(org_v, _field(org_v, fz), None, (bg,)))


def _setup(w=1.0):
    org, phi, wg = _liquid_block(w=w)
    orc = O.OraclePoisson(phi, wg, 1.0, DX)
    coords = _dof_coords(orc)
    assert coords.shape[0] == org.shape[0] * 512
    return orc, wg, coords, _leaf_box(-16, 16)   # velocity leaves cover every face of the block


def test_rhs_uniform_is_divergence_free():
    orc, wg, coords, ov = _setup()
    vx, vy, vz = _vel(ov, *(lambda x, y, z, a=a: np.full(x.shape, a) for a in (0.3, -1.7, 2.5)), bg=0.3)
    sv = _vec_grid(ov, *(lambda x, y, z: np.full(x.shape, 9.0),) * 3, (9.0, 9.0, 9.0))
    assert not orc.build_rhs(wg, vx, vy, vz, sv).any()


def test_rhs_linear_velocity():
    orc, wg, coords, ov = _setup()
    vx, vy, vz = _vel(ov, lambda x, y, z: x * DX, lambda x, y, z: 2 * y * DX, lambda x, y, z: 5 * z * DX)
    b = orc.build_rhs(wg, vx, vy, vz)
    assert np.array_equal(b, np.full_like(b, -8.0))


def test_rhs_quadratic_pins_face_index():
    orc, wg, coords, ov = _setup()
    vx, vy, vz = _vel(ov, lambda x, y, z: x * x * DX, lambda x, y, z: 0 * y, lambda x, y, z: 0 * z)
    b = orc.build_rhs(wg, vx, vy, vz)
    assert np.array_equal(b, -(2.0 * coords[:, 0] + 1.0).astype(np.float32))
    vx, vy, vz = _vel(ov, lambda x, y, z: 0 * x, lambda x, y, z: 0 * y, lambda x, y, z: -z * z * DX)
    b = orc.build_rhs(wg, vx, vy, vz)
    assert np.array_equal(b, (2.0 * coords[:, 2] + 1.0).astype(np.float32))


@pytest.mark.parametrize("w", [0.25, 0.5, 0.75])
def test_rhs_blends_solid_velocity(w):
    orc, wg, coords, ov = _setup(w=w)
    a = (1.0, 2.0, 5.0)     # liquid slopes
    s = (4.0, -3.0, 1.0)    # solid slopes
    vx, vy, vz = _vel(ov, lambda x, y, z: a[0] * x * DX, lambda x, y, z: a[1] * y * DX,
                      lambda x, y, z: a[2] * z * DX)
    sv = _vec_grid(ov, lambda x, y, z: s[0] * x * DX, lambda x, y, z: s[1] * y * DX,
                   lambda x, y, z: s[2] * z * DX, (0.0, 0.0, 0.0))
    b = orc.build_rhs(wg, vx, vy, vz, sv)
    expect = -sum(w * a[c] + (1 - w) * s[c] for c in range(3))
    assert np.array_equal(b, np.full_like(b, expect))
    # without the solid velocity only the liquid part remains
    b0 = orc.build_rhs(wg, vx, vy, vz)
    assert np.array_equal(b0, np.full_like(b0, -w * sum(a)))


def test_rhs_zero_weights_give_zero():
    orc, wg, coords, ov = _setup()
    vx, vy, vz = _vel(ov, lambda x, y, z: x * DX, lambda x, y, z: y * y * DX, lambda x, y, z: 3 * z * DX)
    w0 = _vec_grid(ov, *(lambda x, y, z: np.zeros(x.shape),) * 3, (0.0, 0.0, 0.0))
    sv = _vec_grid(ov, lambda x, y, z: 7 * x * DX, lambda x, y, z: -y * DX, lambda x, y, z: z * z * DX,
                   (1.0, 1.0, 1.0))
    assert not orc.build_rhs(w0, vx, vy, vz, sv).any()


def _brute_rhs(coords, wgrid, vel, solid, dx):
    wm = B.voxel_maps(wgrid.origins, wgrid.values)
    vm = [B.voxel_maps(v.origins, v.values) for v in vel]
    sm = B.voxel_maps(solid.origins, solid.values)
    inv = 1.0 / dx
    out = np.zeros(len(coords))
    for i, g in enumerate(coords):
        g = tuple(int(v) for v in g)
        tot, nz = 0.0, False
        for c in range(3):
            for sgn in (+1, -1):
                f = list(g)
                if sgn > 0:
                    f[c] += 1
                f = tuple(f)
                wt = float(wm[f][0][c]) if f in wm else float(wgrid.background[c])
                u = float(vm[c][f][0]) if f in vm[c] else float(vel[c].background[0])
                us = float(sm[f][0][c]) if f in sm else float(solid.background[c])
                nz |= wt != 0.0
                tot += -sgn * inv * (wt * u + (1.0 - wt) * us)
        out[i] = tot if nz else 0.0
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rhs_matches_dense_brute_force(seed):
    rng = np.random.default_rng(seed)
    org, phi, act, w = B.random_problem(rng, nleaf=5, negative=True)
    wg = G(org, w, None, (1.0, 0.35, 0.0))
    orc = O.OraclePoisson(G(org, phi, act, (0.3,)), wg, 0.37, DX)
    coords = _dof_coords(orc)
    # velocities on a subset of the leaves (the rest read the background), solid velocity on another
    keep_v = rng.random(org.shape[0]) < 0.7
    keep_v[0] = True
    ov = org[keep_v]
    vel = tuple(G(ov, rng.uniform(-1, 1, (ov.shape[0], 512)).astype(np.float32), None, (bg,))
                for bg in (0.5, -0.25, 0.125))
    # one topology for the three velocity channels (the oracle's interface), different backgrounds
    vel = (vel[0], G(ov, vel[1].values, None, (0.5,)), G(ov, vel[2].values, None, (0.5,)))
    keep_s = rng.random(org.shape[0]) < 0.6
    os_ = org[keep_s]
    solid = G(os_, rng.uniform(-2, 2, (os_.shape[0], 512, 3)).astype(np.float32), None, (0.75, -0.5, 0.25))
    got = orc.build_rhs(wg, *vel, solid)
    ref = _brute_rhs(coords, wg, vel, solid, DX)
    assert np.count_nonzero(ref) > 0.5 * ref.size
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
