"""GPU parity of the HIP viscosity path against the CPU oracle (same seeded inputs).

Bit-exact: hierarchy topology, every applier mode (NORMAL, RESIDUAL, JACOBI,
SPAI0, one red / black colour pass), the 6-colour XYZ / ZYX Gauss-Seidel
sweeps, restriction and prolongation. Within 1e-5 relative L2: the coarsest
Eigen-CG restatement (its dot products sum in another order), one W-cycle,
and the PCG result, whose iteration count must match the oracle's within +-1.
A high-viscosity variant (mu = 50) makes the cross-channel terms dominate so a
wrong coupling cannot hide behind the mass term.
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

pytestmark = pytest.mark.gpu

CASES = [(0, 32, 0, 1.0), (1, 32, 1, 1.0), (3, 24, 5, 50.0), (0, 32, 2, 50.0)]


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", params=CASES, ids=[f"cfg{c}-n{n}-s{s}-mu{m:g}" for c, n, s, m in CASES])
def problem(request, hip_available):
    cfg, n, seed, mu = request.param
    _say(f"[setup] viscosity scene cfg={cfg} n={n} seed={seed} mu={mu}")
    s = U.ViscScene(cfg, n, seed=seed, mu=mu)
    lw = U.L_with_level(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx)
    orc = O.OracleViscosity(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx)
    b = orc.build_rhs(s.velocity, None)
    return s, lw, orc, b


def _same(a, b):
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def _rel(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(np.asarray(b, np.float64)), 1e-300))


def test_hierarchy_matches(problem):
    s, lw, orc, b = problem
    assert lw.nlevels == orc.nlevels
    for l in range(lw.nlevels):
        for x, y in zip(lw.level_topology(l), orc.topology(l)):
            assert np.array_equal(x, y), f"level {l}"


def test_rhs_matches(problem):
    s, lw, orc, b = problem
    assert _same(lw.build_rhs(s.velocity).data, b)


def test_apply_modes_bit_exact(problem):
    s, lw, orc, b = problem
    rng = np.random.default_rng(7)
    for l in range(lw.nlevels):
        nd = lw.level_ndof(l)
        x = rng.uniform(-1, 1, nd).astype(np.float32)
        rb = rng.uniform(-1, 1, nd).astype(np.float32)
        for mode in range(8):
            _say(f"  level {l} mode {mode}")
            yg = lw.apply(l, mode, x, rb)
            yo = orc.apply(l, mode, x, rb)
            assert _same(yg, yo), f"level {l} mode {mode}: max diff {np.abs(yg - yo).max()}"


def test_transfers_bit_exact(problem):
    s, lw, orc, b = problem
    rng = np.random.default_rng(11)
    for l in range(lw.nlevels - 1):
        f = rng.uniform(-1, 1, lw.level_ndof(l)).astype(np.float32)
        c = rng.uniform(-1, 1, lw.level_ndof(l + 1)).astype(np.float32)
        assert _same(lw.restriction(l, f), orc.restriction(l, f)), f"restrict {l}"
        assert _same(lw.prolongation(l, c, f), orc.prolongation(l, c, f)), f"prolong {l}"


def test_coarsest_solve(problem):
    s, lw, orc, b = problem
    rng = np.random.default_rng(5)
    rb = rng.uniform(-1, 1, lw.level_ndof(lw.nlevels - 1)).astype(np.float32)
    assert _rel(lw.solve_exact(rb), orc.coarsest_solve(rb)) <= 1e-5


def test_precondition(problem):
    s, lw, orc, b = problem
    solver = U.simd_viscosity3d(lw, s.velocity)
    assert _rel(solver.precondition(b), orc.precondition(b)) <= 1e-5


@pytest.mark.parametrize("tol", [1e-6, 1e-4])
def test_pcg_matches_oracle(problem, tol):
    s, lw, orc, b = problem
    solver = U.simd_viscosity3d(lw, s.velocity)
    x = lw.get_zero_vec()
    solver.pcg_solve(x, tol)
    it_o, xo, _ = orc.solve(b, tol=tol)
    it64, x64 = orc.solve64(b.astype(np.float64), tol=tol)
    _say(f"  tol {tol}: gpu it {solver.m_iteration} | oracle it {it_o}")
    assert abs(solver.m_iteration - it_o) <= 1
    assert _rel(x.data, xo) <= 1e-5
    assert _rel(x.data, x64) <= 1e-5


def test_warm_start_and_max_iterations(problem):
    s, lw, orc, b = problem
    solver = U.simd_viscosity3d(lw, s.velocity)
    # FLIP's warm start: the input velocity on the DOFs
    x0 = b.copy()
    x = x0.copy()
    solver.pcg_solve(x, 1e-6)
    it_o, xo, _ = orc.solve(b, x0=x0, tol=1e-6)
    assert abs(solver.m_iteration - it_o) <= 1
    assert _rel(x, xo) <= 1e-5
    solver.m_max_iter = 1
    x = np.zeros_like(b)
    solver.pcg_solve(x, 1e-12)
    it_o, xo, _ = orc.solve(b, tol=1e-12, maxit=1)
    assert solver.m_iteration == it_o == 0
    assert _rel(x, xo) <= 1e-5


def test_graph_replay_bit_identical(problem):
    """Graph replay and the level-to-kernel mapping reorder launches, not
    arithmetic: flags 2 issues every launch directly, 32 runs coarse levels of
    at most 16 leaves per channel in one workgroup, 128 never does (0 chooses by
    timing)."""
    s, lw, orc, b = problem
    out = []
    for flags in (0, 2, 32, 34, 128):
        l2 = U.L_with_level(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, flags=flags)
        solver = U.simd_viscosity3d(l2, s.velocity)
        x = l2.get_zero_vec()
        solver.pcg_solve(x, 1e-6)
        out.append((x.data.copy(), solver.m_iteration))
    for x, it in out[1:]:
        assert it == out[0][1]
        assert np.array_equal(x, out[0][0])


@pytest.mark.parametrize("env,values", [("UAAMG_VWIDE", ("0", "1")), ("UAAMG_VHALF", ("0", "1", "2"))])
def test_wide_and_part_row_kernels_bit_identical(problem, monkeypatch, env, values):
    """UAAMG_VWIDE=1 runs every level's passes on the two-waves-per-SIMD
    kernels (by default only levels wider than one wave per SIMD), UAAMG_VHALF=1
    on half rows (by default only the other levels), =2 on quarter rows, =0 on
    full rows: a register budget and a lane mapping, not arithmetic, so the
    operator applications and the solve are bit-identical."""
    s, lw, orc, b = problem
    out = []
    for wide in values:
        monkeypatch.setenv(env, wide)
        l2 = U.L_with_level(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx)
        solver = U.simd_viscosity3d(l2, s.velocity)
        x = l2.get_zero_vec()
        solver.pcg_solve(x, 1e-6)
        rng = np.random.default_rng(3)
        xr = rng.standard_normal(l2.level_ndof(0)).astype(np.float32)
        ap = [l2.apply(0, m, xr, b) for m in (0, 1, 2, 4, 5)]
        out.append((x.data.copy(), solver.m_iteration, ap))
    monkeypatch.delenv(env)
    for other in out[1:]:
        assert out[0][1] == other[1]
        assert np.array_equal(out[0][0], other[0])
        for a0, a1 in zip(out[0][2], other[2]):
            assert np.array_equal(a0, a1)
