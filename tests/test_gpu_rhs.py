"""GPU parity of the device right-hand-side build (uaamg_poisson_build_rhs_device):
bit-identical to the oracle's BuildPoissonRhs (simd_vdb_poisson.cpp:19-93) and to the
host build (uaamg_poisson_build_rhs) for host-resident and device-resident grid
values, with a solid velocity and fractional face weights (the solid sphere of the
256^3 scene family), grids in shuffled leaf order, velocity leaves missing
(background lookups with a non-zero background), and the all-zero-weight rule.
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _solid(grid: U.Grid, seed: int) -> U.Grid:
    rng = np.random.default_rng(seed)
    vals = rng.uniform(-1.0, 1.0, size=(grid.origins.shape[0], 512, 3)).astype(np.float32)
    return U.Grid(grid.origins.copy(), vals, None, (0.5, -0.25, 0.125))


def _shuffled(grid: U.Grid, seed: int) -> U.Grid:
    p = np.random.default_rng(seed).permutation(grid.origins.shape[0])
    act = None if grid.active is None else grid.active[p].copy()
    return U.Grid(grid.origins[p].copy(), grid.values[p].copy(), act, grid.background)


def _thinned(grid: U.Grid, seed: int, bg: float) -> U.Grid:
    keep = np.random.default_rng(seed).uniform(size=grid.origins.shape[0]) > 0.2
    act = None if grid.active is None else grid.active[keep].copy()
    return U.Grid(grid.origins[keep].copy(), grid.values[keep].copy(), act, (bg, bg, bg))


@pytest.mark.parametrize("cfg,n,seed", [(0, 32, 0), (1, 48, 1), (3, 40, 3)])
def test_device_rhs_equals_host(hip_available, cfg, n, seed):
    import torch
    dev = torch.device("cuda:0")
    _say(f"[setup] rhs cfg={cfg} n={n} seed={seed}")
    s = U.Scene(cfg, n, seed=seed)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
    solid = _solid(s.weights, seed + 7)
    cases = [
        ("plain", (s.weights, s.vx, s.vy, s.vz, None)),
        ("solid velocity", (s.weights, s.vx, s.vy, s.vz, solid)),
        ("shuffled leaves, missing velocity leaves", (_shuffled(s.weights, seed), _thinned(s.vx, seed, 0.25),
                                                     _shuffled(s.vy, seed + 1), _thinned(s.vz, seed + 2, -0.5),
                                                     _shuffled(solid, seed + 3))),
    ]
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    if cfg == 1:
        frac = s.weights.values[(s.weights.values > 0) & (s.weights.values < 1)].size
        assert frac > 0, "the solid sphere gives fractional face weights"
    for name, (w, vx, vy, vz, sv) in cases:
        ref = lap.createPressurePoissonRightHandSide(w, vx, vy, vz, sv)
        if not name.startswith("shuffled"):
            # the oracle takes the three velocity channels on one topology (the scene's)
            assert np.array_equal(ref, orc.build_rhs(w, vx, vy, vz, sv)), f"{name}: host build vs oracle"
        out = torch.full((lap.mNumDof,), float("nan"), device=dev)
        lap.createPressurePoissonRightHandSide_device(out.data_ptr(), w, vx, vy, vz, sv)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        _say(f"  {name}: ndof {lap.mNumDof}, nonzero {int(np.count_nonzero(ref))}")
        assert np.array_equal(got, ref), f"{name}: max diff {np.abs(got - ref).max()}"
        if not name.startswith("shuffled"):
            assert np.array_equal(got, orc.build_rhs(w, vx, vy, vz, sv)), f"{name}: device build vs oracle"
        # the same grids with their values already in device memory
        grids = [w, vx, vy, vz] + ([] if sv is None else [sv])
        dvals = [torch.from_numpy(np.ascontiguousarray(g.values, dtype=np.float32)).to(dev) for g in grids]
        out.fill_(float("nan"))
        lap.createPressurePoissonRightHandSide_device(out.data_ptr(), w, vx, vy, vz, sv,
                                                      values_ptrs=[t.data_ptr() for t in dvals])
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref), f"{name} (device values)"


def test_device_rhs_zero_weights(hip_available):
    """all face weights 0: every DOF's rhs is 0 (the hasNonZeroWeight rule), whatever the velocities."""
    import torch
    s = U.Scene(0, 32, seed=2)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
    w0 = U.Grid(s.weights.origins.copy(), np.zeros_like(s.weights.values), None, (0.0, 0.0, 0.0))
    out = torch.full((lap.mNumDof,), float("nan"), device=torch.device("cuda:0"))
    lap.createPressurePoissonRightHandSide_device(out.data_ptr(), w0, s.vx, s.vy, s.vz, _solid(s.weights, 3))
    torch.cuda.synchronize()
    assert not out.cpu().numpy().any()
    assert not lap.createPressurePoissonRightHandSide(w0, s.vx, s.vy, s.vz, _solid(s.weights, 3)).any()
