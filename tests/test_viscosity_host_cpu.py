"""CPU-side checks of the product's viscosity host build (no GPU needed).

The product builds the viscosity hierarchy with its own hash-map / leaf-block
code (uaamg_amd/csrc/viscosity_host.cpp); a device=-1 handle exposes it for
inspection. It must equal the oracle bit for bit: per-level DOF leaves, masks
and numbering, every coefficient grid with its kept / trimmed leaves (SPAI0
included), the explicit matrices and the right-hand side (with and without a
solid velocity). Solves on a device=-1 handle fail loudly.
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

CASES = [(0, 32, 0, 1.0, False), (1, 32, 1, 1.0, True), (3, 24, 5, 50.0, True), (1, 40, 3, 5.0, False)]


@pytest.mark.parametrize("cfg,n,seed,mu,solid_vel", CASES)
def test_host_hierarchy_bit_exact_vs_oracle(cfg, n, seed, mu, solid_vel):
    s = U.ViscScene(cfg, n, seed=seed, mu=mu)
    lw = U.L_with_level(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, device=-1)
    orc = O.OracleViscosity(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx)
    assert lw.nlevels == orc.nlevels
    for l in range(lw.nlevels):
        for a, b in zip(lw.level_topology(l), orc.topology(l)):
            assert np.array_equal(a, b), f"level {l} topology"
        assert lw.level_ndof(l) == orc.level_ndof(l)
        for gid in range(30):
            pc, pk = lw.level_coefficients(l, gid)
            oc, ok = orc.coefficients(l, gid)
            assert np.array_equal(pk, ok), f"level {l} grid {gid}: kept leaves differ"
            assert np.array_equal(pc, oc), f"level {l} grid {gid}: values differ"
        for a, b in zip(lw.level_matrix(l), orc.matrix(l)):
            assert np.array_equal(a, b), f"level {l} explicit matrix"
    us = None
    if solid_vel:
        rng = np.random.default_rng(seed)
        us = U.Grid(s.velocity.origins, rng.uniform(-1, 1, s.velocity.values.shape).astype(np.float32), None,
                    (0.3, -0.2, 0.1))
    rp = lw.build_rhs(s.velocity, us).data
    ro = orc.build_rhs(s.velocity, us)
    assert np.array_equal(rp, ro)


def test_untrimmed_build_matches():
    s = U.ViscScene(0, 24, seed=4, mu=3.0)
    lw = U.L_with_level(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, device=-1, flags=8)
    orc = O.OracleViscosity(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, trim=False)
    for gid in range(27):
        pc, pk = lw.level_coefficients(0, gid)
        oc, ok = orc.coefficients(0, gid)
        assert np.array_equal(pk, ok) and np.array_equal(pc, oc), gid


def test_dof_numbering_is_channel_major_tree_order():
    s = U.ViscScene(1, 24, seed=2)
    lw = U.L_with_level(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, device=-1)
    org, mask, base = lw.level_topology(0)
    nchan = lw.level_info(0)["nchan"]
    start = [0, nchan[0], nchan[0] + nchan[1]]
    for c in range(3):
        counts = np.array([sum(bin(int(w)).count("1") for w in mask[i, c]) for i in range(org.shape[0])])
        has = counts > 0
        # DOFs of one channel are consecutive over its leaves in tree order
        assert base[has, c][0] == start[c]
        assert np.array_equal(base[has, c][1:], (base[has, c] + counts[has])[:-1])
    assert sum(nchan) == lw.m_ndof


def test_host_only_handle_refuses_solves():
    s = U.ViscScene(0, 16, seed=0)
    lw = U.L_with_level(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, device=-1)
    solver = U.simd_viscosity3d(lw, s.velocity)
    x = lw.get_zero_vec()
    with pytest.raises(U.UaamgError):
        solver.pcg_solve(x, 1e-6)
    with pytest.raises(U.UaamgError):
        lw.L_apply(x.data)


def test_bad_inputs_are_rejected():
    s = U.ViscScene(0, 16, seed=0)
    with pytest.raises(U.UaamgError):
        U.L_with_level(s.viscosity, s.liquid, s.solid, -1.0, s.rho, s.dx, device=-1)
    bad = U.Grid(s.liquid.origins + 1, s.liquid.values, s.liquid.active, s.liquid.background)
    with pytest.raises(U.UaamgError):
        U.L_with_level(s.viscosity, bad, s.solid, s.dt, s.rho, s.dx, device=-1)
    with pytest.raises(U.UaamgError):
        U.L_with_level(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, device=-1, max_coarsest_dof=5000)


def test_visc_scene_closed_forms():
    s = U.ViscScene(0, 32, seed=0)
    n, dx = 32, s.dx
    # container SDF on the vertex lattice: dx * distance to the nearest wall index (0 or n)
    org, val = s.solid.origins, s.solid.values
    off = np.arange(512)
    for i in range(0, org.shape[0], 7):
        x = org[i, 0] + (off >> 6)
        y = org[i, 1] + ((off >> 3) & 7)
        z = org[i, 2] + (off & 7)
        m = np.minimum.reduce([x, n - x, y, n - y, z, n - z]).astype(np.float32) * np.float32(dx)
        want = np.where(np.abs(m) < 3 * dx, m, np.where(m < 0, -3 * dx, 3 * dx)).astype(np.float32)
        assert np.array_equal(val[i], want)
    assert s.viscosity.origins.shape[0] == 0 and s.viscosity.background[0] == 1.0
    v = s.velocity.values
    assert v.shape[-1] == 3 and np.all(np.abs(v) <= 1.0) and abs(float(v.mean())) < 0.05
