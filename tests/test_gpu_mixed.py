"""GPU checks of the mixed-precision Poisson form (BF16-stored W-cycle vectors,
FP32 arithmetic, FP64 dot products, FP64 residual replacement).

The reference has no such mode, so the oracle fixes what must hold instead of
bit patterns: the solution agrees with the FP64 oracle solve within 1e-5
relative L2; the FP64 true residual ||b - A x||_inf, recomputed here on the
host from the oracle's FP64 operator, meets the 1e-7 relative tolerance (the
FP32 form only guarantees that for its recursive residual); the W-cycle is
within BF16 rounding of the FP32 oracle's; graph replay reorders launches but
not arithmetic (bit-identical).
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

pytestmark = pytest.mark.gpu

CASES = [(0, 32, 0), (1, 48, 1), (3, 40, 5), (1, 96, 2)]
TOL = 1e-7
BF16 = U.PRECISION_BF16_CYCLE


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", params=CASES, ids=[f"cfg{c}-n{n}-s{s}" for c, n, s in CASES])
def problem(request, hip_available):
    cfg, n, seed = request.param
    _say(f"[setup] mixed-precision scene cfg={cfg} n={n} seed={seed}")
    s = U.Scene(cfg, n, seed=seed)
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    b = orc.build_rhs(s.weights, s.vx, s.vy, s.vz)
    st64, it64, x64 = orc.solve_pcg64(b.astype(np.float64))
    assert st64 == 0
    return s, orc, b, x64


def _rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _solve(s, b, precision, flags=0, tol=TOL, x0=None):
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, flags=flags, precision=precision)
    solver = U.PoissonSolver(lap)
    solver.mRelativeTolerance = tol
    x = lap.getZeroVectorGrid() if x0 is None else x0.copy()
    st = solver.solveMultigridPCG(x, b)
    return lap, solver, st, x


@pytest.mark.parametrize("flags", [0, 16])
def test_mixed_solution_and_true_residual(problem, flags):
    """flags 16 runs the coarse levels through the per-sweep kernels instead of
    the block-visit kernels (both store BF16, rounding at the same points)."""
    s, orc, b, x64 = problem
    lap, solver, st, x = _solve(s, b, BF16, flags)
    _, solver32, st32, x32 = _solve(s, b, U.PRECISION_FP32)
    true_nu, reps = lap.last_residual()
    r = b.astype(np.float64) - orc.apply64(x.astype(np.float64))
    nu_host = float(np.abs(r).max())
    nu0 = float(np.abs(b.astype(np.float64)).max())   # zero initial guess: r0 = b
    _say(f"  flags {flags}: bf16 it {solver.mIterationTaken} (replacements {reps}) | fp32 it "
         f"{solver32.mIterationTaken}; true rel residual {nu_host / nu0:.3e}; err vs fp64 {_rel(x, x64):.3e}")
    assert st == U.SUCCESS and st32 == U.SUCCESS
    assert reps >= 0
    assert nu_host <= 1.01 * TOL * nu0
    assert abs(true_nu - nu_host) <= 1e-3 * nu_host + 1e-30
    assert _rel(x, x64) <= 1e-5
    # BF16 storage perturbs the preconditioner slightly; it must not cost
    # more than a few iterations
    assert solver.mIterationTaken <= solver32.mIterationTaken + 3


def test_mixed_w_cycle_within_bf16_rounding(problem):
    s, orc, b, x64 = problem
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, precision=BF16)
    z = U.PoissonSolver(lap).precondition(b)
    zo = orc.precondition(b, 0)
    err = _rel(z, zo)
    _say(f"  W-cycle rel L2 vs FP32 oracle: {err:.3e}")
    # each stored value carries a relative rounding error <= 2^-9; the cycle
    # composes a few dozen of them
    assert err <= 3e-2


def test_mixed_graph_replay_and_kernel_mapping_bit_identical(problem):
    """flags 2: no graph replay; 16: per-sweep kernels on every coarse level."""
    s, orc, b, x64 = problem
    out = []
    for flags in (0, 2, 16, 18):
        lap, solver, st, x = _solve(s, b, BF16, flags)
        out.append((st, solver.mIterationTaken, x))
    for st, it, x in out[1:]:
        assert st == out[0][0] and it == out[0][1]
        assert np.array_equal(x, out[0][2])


def test_fp32_with_residual_replacement(problem):
    """flags 64: the reference's FP32 arithmetic plus the FP64 replacement step."""
    s, orc, b, x64 = problem
    lap, solver, st, x = _solve(s, b, U.PRECISION_FP32, 64)
    r = b.astype(np.float64) - orc.apply64(x.astype(np.float64))
    nu0 = float(np.abs(b.astype(np.float64)).max())
    assert st == U.SUCCESS
    assert float(np.abs(r).max()) <= 1.01 * TOL * nu0
    assert _rel(x, x64) <= 1e-5
    # without the flag no replacement is reported
    lap2, solver2, st2, x2 = _solve(s, b, U.PRECISION_FP32, 0)
    assert lap2.last_residual() == (-1.0, -1)


def test_mixed_warm_start_and_failure(problem):
    s, orc, b, x64 = problem
    # warm start from the solution perturbed by 5% (the serving stream's form)
    rng = np.random.default_rng(13)
    x0 = (x64 * (1.0 + 0.05 * rng.uniform(-1, 1, x64.shape))).astype(np.float32)
    lap, solver, st, x = _solve(s, b, BF16, x0=x0)
    assert st == U.SUCCESS
    assert _rel(x, x64) <= 1e-5
    # one iteration cannot reach 1e-7
    solver.mMaxIteration = 1
    x3 = lap.getZeroVectorGrid()
    assert solver.solveMultigridPCG(x3, b) == U.FAILED
    # a NaN in the right-hand side fails, as in the FP32 form
    solver.mMaxIteration = 6
    bb = b.copy()
    bb[len(bb) // 2] = np.nan
    x4 = lap.getZeroVectorGrid()
    assert solver.solveMultigridPCG(x4, bb) == U.FAILED


def test_mixed_warm_started_stream_of_right_hand_sides(problem):
    """A serving stream: each solve starts from the previous right-hand
    side's solution (the residual is then mostly the 5% perturbation, so the
    tolerance lies near the FP32 iterate's floor). Every solve succeeds, and
    the FP64 true residual it reports is the host's."""
    s, orc, b, x64 = problem
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, precision=BF16)
    solver = U.PoissonSolver(lap)
    solver.mRelativeTolerance = TOL
    rng = np.random.default_rng(21)
    x = lap.getZeroVectorGrid()
    for k in range(6):
        bk = (b * (1.0 + 0.05 * rng.standard_normal(b.shape))).astype(np.float32)
        st = solver.solveMultigridPCG(x, bk)
        true_nu, reps = lap.last_residual()
        r = bk.astype(np.float64) - orc.apply64(x.astype(np.float64))
        nu_host = float(np.abs(r).max())
        _say(f"  warm solve {k}: status {st}, it {solver.mIterationTaken}, replacements {reps}, "
             f"true residual {nu_host:.3e}")
        assert st == U.SUCCESS, f"solve {k}"
        assert abs(true_nu - nu_host) <= 1e-3 * nu_host + 1e-30


def test_mixed_rejects_unknown_precision(problem):
    s, orc, b, x64 = problem
    with pytest.raises(U.UaamgError):
        U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, precision=7)
