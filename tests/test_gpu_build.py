"""GPU parity of the device stage of the level-0 build (finest_device.h): a handle
built on the device has exactly the hierarchy of a host-only handle (device -1) -
level sizes, DOF topology, every coefficient brick and trim flag on every level.
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
