"""GPU parity of the device hierarchy build (hierarchy_device.h: the level-0
per-voxel stage, compaction, trimming and 1/diag, the coarse Galerkin sums and
the packing into the solver's arrays, all on the device): a handle built on the
device has exactly the hierarchy of a host-only handle (device -1) - level
sizes, DOF topology, every coefficient brick and trim flag on every level.
"""
import numpy as np
import pytest

import uaamg_amd as U

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,n,seed", [(0, 32, 0), (1, 64, 1), (3, 48, 3)])
def test_device_build_equals_host_build(hip_available, cfg, n, seed):
    print(f"[case] build cfg={cfg} n={n} seed={seed}", flush=True)
    s = U.Scene(cfg, n, seed=seed)
    dev = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=0)
    host = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
    assert dev.nlevels == host.nlevels and dev.mNumDof == host.mNumDof
    for level in range(dev.nlevels):
        assert dev.level_info(level) == host.level_info(level)
        for a, b in zip(dev.level_topology(level), host.level_topology(level)):
            assert np.array_equal(a, b), f"level {level} topology"
        for kind in range(4):
            (cd, kd), (ch, kh) = dev.level_coefficients(level, kind), host.level_coefficients(level, kind)
            assert np.array_equal(kd, kh), f"level {level} kind {kind} trim flags"
            assert np.array_equal(cd.view(np.uint32), ch.view(np.uint32)), f"level {level} kind {kind} bricks"


@pytest.mark.parametrize("cfg,n,seed", [(1, 128, 0)])
def test_device_resident_hierarchy_at_128(hip_available, cfg, n, seed):
    """The hierarchy built and kept on the device (compaction, trimming, Galerkin
    sums, packing) against the host build at 128^3: every level, every brick and
    flag, and the level-0 explicit matrix."""
    print(f"[case] device-resident build cfg={cfg} n={n} seed={seed}", flush=True)
    s = U.Scene(cfg, n, seed=seed)
    dev = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=0)
    host = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
    assert dev.nlevels == host.nlevels and dev.mNumDof == host.mNumDof
    for level in range(dev.nlevels):
        assert dev.level_info(level) == host.level_info(level)
        for a, b in zip(dev.level_topology(level), host.level_topology(level)):
            assert np.array_equal(a, b), f"level {level} topology"
        for kind in range(5):
            (cd, kd), (ch, kh) = dev.level_coefficients(level, kind), host.level_coefficients(level, kind)
            assert np.array_equal(kd, kh), f"level {level} kind {kind} trim flags"
            assert np.array_equal(cd.view(np.uint32), ch.view(np.uint32)), f"level {level} kind {kind} bricks"
    for a, b in zip(dev.getTriplets(0), host.getTriplets(0)):
        assert np.array_equal(a, b)


def test_device_resident_spai0_at_creation(hip_available):
    """A handle created with the SPAI0 smoother copies the device-built bricks to the
    host at creation (for the SPAI0 build); its bricks and its solve match a handle
    whose bricks stayed on the host."""
    print("[case] device-resident build with SPAI0 at creation", flush=True)
    s = U.Scene(3, 48, seed=5)
    dev = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=0, smoother=3)
    ref = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=0, smoother=3, flags=256)
    for level in range(dev.nlevels):
        for kind in range(5):
            (cd, kd), (ch, kh) = dev.level_coefficients(level, kind), ref.level_coefficients(level, kind)
            assert np.array_equal(kd, kh) and np.array_equal(cd.view(np.uint32), ch.view(np.uint32))
    b = np.random.default_rng(5).standard_normal(dev.mNumDof).astype(np.float32)
    outs = []
    for lap in (dev, ref):
        sol = U.PoissonSolver(lap)
        sol.mSmoother = 3
        x = np.zeros_like(b)
        status = sol.solveMultigridPCG(x, b)
        outs.append((status, sol.mIterationTaken, x))
    assert outs[0][0] == outs[1][0] and outs[0][1] == outs[1][1]
    assert np.array_equal(outs[0][2].view(np.uint32), outs[1][2].view(np.uint32))
