"""The viscosity level-0 volume stage on the device (subcell SDF, subcell
volumes, voxel / face / edge volumes; simd_viscosity3d.cpp:258-498) against the
host stage (flags bit 8): every level's topology, DOF numbering and all 30
coefficient grids, and the right-hand side, bit for bit, on scenes with pools,
spheres and solids, including a solid grid with missing leaves' background."""
import numpy as np
import pytest

from uaamg_amd.viscosity import L_with_level, ViscScene

pytestmark = pytest.mark.gpu

CASES = [(0, 24, 0, 1.0), (3, 40, 5, 1.0), (0, 64, 1, 20.0), (3, 96, 7, 1.0)]


@pytest.mark.parametrize("cfg,n,seed,mu", CASES)
def test_device_volume_stage_equals_host(hip_available, cfg, n, seed, mu):
    print(f"[case] viscosity build cfg={cfg} n={n} seed={seed}", flush=True)
    v = ViscScene(cfg, n, seed=seed, mu=mu)
    dev = L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx, device=0)
    host = L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx, device=0, flags=256)
    assert dev.nlevels == host.nlevels and dev.m_ndof == host.m_ndof
    for lv in range(dev.nlevels):
        for a, b in zip(dev.level_topology(lv), host.level_topology(lv)):
            assert np.array_equal(a, b), f"level {lv} topology"
        for g in range(27):
            ca, ka = dev.level_coefficients(lv, g)
            cb, kb = host.level_coefficients(lv, g)
            assert np.array_equal(ka, kb) and np.array_equal(ca, cb), f"level {lv} grid {g}"
    assert np.array_equal(dev.build_rhs(v.velocity).data, host.build_rhs(v.velocity).data)
    print(f"  ndof {dev.m_ndof}, levels {dev.nlevels}: identical", flush=True)
