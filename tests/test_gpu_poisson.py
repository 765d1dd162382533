"""GPU parity of the HIP Poisson path against the CPU oracle (same seeded inputs).

Bit-exact: hierarchy topology and DOF numbering, every operator mode of the
7-point apply (Laplacian, residual, Jacobi, SPAI0, red/black GS), first sweeps,
restriction and prolongation. Within 1e-5 relative L2: the coarsest Eigen-CG
restatement (its dot products sum in a different order), one W-cycle, and the
PCG solution, whose iteration count must match the oracle's within +-1.
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

pytestmark = pytest.mark.gpu

CASES = [(0, 32, 0), (1, 48, 1), (3, 40, 5)]


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", params=CASES, ids=[f"cfg{c}-n{n}-s{s}" for c, n, s in CASES])
def problem(request, hip_available):
    cfg, n, seed = request.param
    _say(f"[setup] scene cfg={cfg} n={n} seed={seed}")
    s = U.Scene(cfg, n, seed=seed)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    b = orc.build_rhs(s.weights, s.vx, s.vy, s.vz)
    return s, lap, orc, b


def _same(a, b):
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def _rel(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b.astype(np.float64)), 1e-300))


def test_hierarchy_matches(problem):
    s, lap, orc, b = problem
    assert lap.nlevels == orc.nlevels
    for l in range(lap.nlevels):
        po, pm, pb = lap.level_topology(l)
        oo, om, ob = orc.topology(l)
        assert np.array_equal(po, oo) and np.array_equal(pm, om) and np.array_equal(pb, ob), f"level {l}"
        assert lap.level_info(l)["nleaf_tree"] == orc.nleaf_tree(l)


def test_apply_modes_bit_exact(problem):
    s, lap, orc, b = problem
    rng = np.random.default_rng(7)
    for l in range(lap.nlevels):
        nd = lap.level_ndof(l)
        x = rng.uniform(-1, 1, nd).astype(np.float32)
        rb = rng.uniform(-1, 1, nd).astype(np.float32)
        for mode, w in [(0, 0.0), (1, 0.0), (2, 2.2473), (2, 0.5296), (3, 0.0), (4, 1.2), (5, 1.2), (4, 1.0)]:
            _say(f"  level {l} mode {mode} w {w}")
            yg = lap.apply(l, mode, x, rb, w)
            yo = orc.apply(l, mode, x, rb, w)
            assert _same(yg, yo), f"level {l} mode {mode}: max diff {np.abs(yg - yo).max()}"


def test_first_sweeps_bit_exact(problem):
    s, lap, orc, b = problem
    rng = np.random.default_rng(3)
    for l in range(lap.nlevels):
        rb = rng.uniform(-1, 1, lap.level_ndof(l)).astype(np.float32)
        for kind, w in [(0, 2.2473), (0, 6 / 7), (1, 0.0), (2, 1.2)]:
            yg = lap.first_sweep(l, kind, rb, w)
            yo = orc.first_sweep(l, kind, rb, w)
            assert _same(yg, yo), f"level {l} kind {kind}"


def test_transfers_bit_exact(problem):
    s, lap, orc, b = problem
    rng = np.random.default_rng(11)
    for l in range(lap.nlevels - 1):
        f = rng.uniform(-1, 1, lap.level_ndof(l)).astype(np.float32)
        c = rng.uniform(-1, 1, lap.level_ndof(l + 1)).astype(np.float32)
        assert _same(lap.restriction(l, f), orc.restriction(l, f)), f"restrict {l}"
        for alpha in (1.0, 0.5):
            assert _same(lap.prolongation(l, c, f, alpha), orc.prolongation(l, c, f, alpha)), f"prolong {l}"


def test_coarsest_solve(problem):
    s, lap, orc, b = problem
    rng = np.random.default_rng(5)
    rb = rng.uniform(-1, 1, lap.level_ndof(lap.nlevels - 1)).astype(np.float32)
    xg = lap.coarsest_solve(rb)
    xo = orc.coarsest_solve(rb)
    assert _rel(xg, xo) <= 1e-5


@pytest.mark.parametrize("smoother", [0, 1, 2, 3])
def test_precondition(problem, smoother):
    s, lap, orc, b = problem
    solver = U.PoissonSolver(lap)
    solver.mSmoother = smoother
    zg = solver.precondition(b)
    zo = orc.precondition(b, smoother)
    assert _rel(zg, zo) <= 1e-5


@pytest.mark.parametrize("smoother", [0, 1, 2, 3])
def test_pcg_matches_oracle(problem, smoother):
    s, lap, orc, b = problem
    solver = U.PoissonSolver(lap)
    solver.mSmoother = smoother
    x = lap.getZeroVectorGrid()
    st = solver.solveMultigridPCG(x, b)
    st_o, it_o, xo = orc.solve_pcg(b, smoother=smoother)
    st64, it64, x64 = orc.solve_pcg64(b.astype(np.float64), smoother=smoother)
    _say(f"  smoother {smoother}: gpu it {solver.mIterationTaken} st {st} | oracle it {it_o} st {st_o}")
    assert st == U.SUCCESS and st_o == 0
    assert abs(solver.mIterationTaken - it_o) <= 1
    assert _rel(x, xo) <= 1e-5
    assert _rel(x, x64) <= 1e-5


def test_pure_multigrid(problem):
    s, lap, orc, b = problem
    solver = U.PoissonSolver(lap)
    solver.mMaxIteration = 4
    x = lap.getZeroVectorGrid()
    st = solver.solvePureMultigrid(x, b)
    st_o, it_o, xo = orc.solve_mg(b, maxit=4)
    assert st == st_o
    assert abs(solver.mIterationTaken - it_o) <= 1
    assert _rel(x, xo) <= 1e-5


def test_warm_start_and_tolerance(problem):
    s, lap, orc, b = problem
    solver = U.PoissonSolver(lap)
    solver.mRelativeTolerance = 1e-4
    x = lap.getZeroVectorGrid()
    assert solver.solveMultigridPCG(x, b) == U.SUCCESS
    it_loose = solver.mIterationTaken
    st_o, it_o, xo = orc.solve_pcg(b, tol=1e-4)
    assert abs(it_loose - it_o) <= 1
    # already converged: a second solve from the solution returns at once
    solver.mRelativeTolerance = 1e-7
    x2 = x.copy()
    st2 = solver.solveMultigridPCG(x2, b)
    st_o2, it_o2, _ = orc.solve_pcg(b, x0=x, tol=1e-7)
    assert st2 == st_o2 == 0
    assert abs(solver.mIterationTaken - it_o2) <= 1


def test_failed_status_on_max_iterations(problem):
    s, lap, orc, b = problem
    solver = U.PoissonSolver(lap)
    solver.mMaxIteration = 1
    x = lap.getZeroVectorGrid()
    assert solver.solveMultigridPCG(x, b) == U.FAILED
    st_o, it_o, _ = orc.solve_pcg(b, maxit=1)
    assert st_o == 1 and solver.mIterationTaken == it_o == 1


def test_nan_rhs_fails(problem):
    s, lap, orc, b = problem
    solver = U.PoissonSolver(lap)
    solver.mMaxIteration = 6
    bb = b.copy()
    bb[len(bb) // 2] = np.nan
    x = lap.getZeroVectorGrid()
    assert solver.solveMultigridPCG(x, bb) == U.FAILED


@pytest.mark.parametrize("smoother", [0, 1, 2, 3])
def test_fused_and_graph_paths_bit_identical(problem, smoother):
    """Fused kernels, the block-visit kernels, the bottom-levels tail kernel and
    graph replay reorder launches, not arithmetic: flags 1 = unfused, 2 = no
    graphs, 4 = bottom levels in the single-workgroup tail kernel, 16 = no
    block-visit kernels on the coarse levels."""
    s, lap, orc, b = problem
    variants = {}
    for flags in (0, 1, 2, 3, 4, 6, 16, 18, 20):
        l2 = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, flags=flags)
        solver = U.PoissonSolver(l2)
        solver.mSmoother = smoother
        z = solver.precondition(b)
        x = l2.getZeroVectorGrid()
        st = solver.solveMultigridPCG(x, b)
        variants[flags] = (z, x, st, solver.mIterationTaken)
    z0, x0, st0, it0 = variants[0]
    for flags, (z, x, st, it) in variants.items():
        assert np.array_equal(z, z0), f"flags {flags}: W-cycle differs"
        assert st == st0 and it == it0, f"flags {flags}: {st}/{it} vs {st0}/{it0}"
        assert np.array_equal(x, x0), f"flags {flags}: solution differs"


@pytest.mark.parametrize("precision", [U.PRECISION_FP32, U.PRECISION_BF16_CYCLE])
def test_coarsest_inside_block_post_visit_bit_identical(problem, monkeypatch, precision):
    """With block-visit kernels on every coarse level (UAAMG_BLOCK_LEAVES), the post-visit of the
    level above the coarsest runs the coarsest CG in its own launch, each workgroup solving the
    whole coarsest system from the same routine; UAAMG_POST_CG=0 launches the CG on its own. The
    W-cycle, the PCG solution and its iteration count are bit-identical either way, and the
    kernel map reports the fused form only where it runs."""
    s, lap, orc, b = problem
    monkeypatch.setenv("UAAMG_BLOCK_LEAVES", "1000000")
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("UAAMG_POST_CG", fused)
        l2 = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, precision=precision)
        km = l2.kernel_map()
        L = l2.nlevels
        assert [k["post_cg"] for k in km] == [fused == "1" and L >= 3 and l == L - 2 for l in range(L)]
        solver = U.PoissonSolver(l2)
        z = solver.precondition(b)
        x = l2.getZeroVectorGrid()
        st = solver.solveMultigridPCG(x, b)
        out[fused] = (z, x, st, solver.mIterationTaken)
        _say(f"  precision {precision}, post_cg {fused}: levels {L}, status {st}, iterations {solver.mIterationTaken}")
    (z1, x1, s1, i1), (z0, x0, s0, i0) = out["1"], out["0"]
    assert np.array_equal(z1, z0)
    assert (s1, i1) == (s0, i0) and s1 == U.SUCCESS
    assert np.array_equal(x1, x0)


def test_coarsest_of_more_than_2048_rows_inside_block_post_visit(hip_available, monkeypatch):
    """The 4096-row forms of the coarsest CG (standalone and inside the block post-visit): a 72^3
    pool scene whose three-level hierarchy ends in a 2736-row coarsest system, with block-visit
    kernels forced on level 1. Fused and separate solves are bit-identical in both precisions, and
    the FP32 solve matches the oracle (status, iterations within 1, 1e-5 relative L2 against it and
    its FP64 twin)."""
    _say("[case] Scene(0, 72, seed=0): coarsest CG of more than 2048 rows, fused and separate")
    s = U.Scene(0, 72, seed=0)
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    b = orc.build_rhs(s.weights, s.vx, s.vy, s.vz)
    monkeypatch.setenv("UAAMG_BLOCK_LEAVES", "1000000")
    for precision in (U.PRECISION_FP32, U.PRECISION_BF16_CYCLE):
        out = {}
        for fused in ("1", "0"):
            monkeypatch.setenv("UAAMG_POST_CG", fused)
            lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, precision=precision)
            L = lap.nlevels
            assert L == 3 and 2048 < lap.level_info(L - 1)["ndof"] <= 4096
            assert lap.kernel_map()[1]["post_cg"] == (fused == "1")
            solver = U.PoissonSolver(lap)
            z = solver.precondition(b)
            x = lap.getZeroVectorGrid()
            st = solver.solveMultigridPCG(x, b)
            out[fused] = (z, x, st, solver.mIterationTaken)
        (z1, x1, s1, i1), (z0, x0, s0, i0) = out["1"], out["0"]
        _say(f"  precision {precision}: status {s1}, iterations {i1}")
        assert np.array_equal(z1, z0) and np.array_equal(x1, x0) and (s1, i1) == (s0, i0)
        assert s1 == U.SUCCESS
        if precision == U.PRECISION_FP32:
            st_o, it_o, xo = orc.solve_pcg(b)
            _, _, x64 = orc.solve_pcg64(b.astype(np.float64))
            assert st_o == 0 and abs(i1 - it_o) <= 1
            assert _rel(x1, xo) <= 1e-5 and _rel(x1, x64) <= 1e-5
