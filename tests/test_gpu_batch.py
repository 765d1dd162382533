"""GPU parity of the merged batch (uaamg_poisson_batch_*): every problem's
status, iteration count (within 1) and solution (1e-5 relative L2) agree with
the oracle's FP32 solveMultigridPCG and with its FP64 twin; every problem's
result, status and iteration count are bit-identical to solving it alone with
its own handle, for a batch of uneven problems (one-, two- and three-level
hierarchies, given out of depth order), in FP32 and in the BF16 W-cycle form,
with and without graph replay; per-problem failure (iteration limit, a NaN
right-hand side) and a zero right-hand side stay confined to their problem;
FLIP's pure-multigrid fallback (FLIP_vdb.cpp:3150-3155) reproduces the
oracle's solvePureMultigrid on the problems whose PCG fails, including the
seeded 36^3 problem on which the reference's own PCG rule fails.
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

pytestmark = pytest.mark.gpu

# (config, n, seed): hierarchies of 2, 1, 3, 2, 3, 1, 2 levels
PROBLEMS = [(3, 24, 1), (3, 16, 0), (1, 48, 1), (3, 40, 3), (3, 64, 7), (3, 16, 3), (0, 32, 0)]


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def batch_inputs(hip_available):
    import torch
    dev = torch.device("cuda:0")
    scenes, rhs = [], []
    for cfg, n, seed in PROBLEMS:
        _say(f"[setup] batch problem cfg={cfg} n={n} seed={seed}")
        s = U.Scene(cfg, n, seed=seed)
        lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
        scenes.append(s)
        rhs.append(torch.from_numpy(lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)).to(dev))
    return scenes, rhs


def _alone(scenes, rhs, precision=0, flags=0, max_iterations=100):
    import torch
    out = []
    for s, b in zip(scenes, rhs):
        lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, precision=precision, flags=flags)
        solver = U.PoissonSolver(lap)
        solver.mMaxIteration = max_iterations
        solver._push_options()
        x = torch.zeros_like(b)
        rc, it = lap.solve_pcg_device(x.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        out.append((rc, it, x.cpu().numpy(), lap.last_residual()))
    return out


def _batched(scenes, rhs, precision=0, flags=0, max_iterations=100, x0=None):
    import torch
    bt = U.PoissonBatch([(s.phi, s.weights, 1.0, s.dx) for s in scenes], precision=precision, flags=flags,
                        max_iterations=max_iterations)
    xs = [torch.zeros_like(b) if x0 is None else x0[i].clone() for i, b in enumerate(rhs)]
    st, it = bt.solve_device([x.data_ptr() for x in xs], [b.data_ptr() for b in rhs])
    torch.cuda.synchronize()
    return bt, [(st[i], it[i], xs[i].cpu().numpy()) for i in range(len(rhs))]


@pytest.mark.parametrize("precision,flags", [(0, 0), (0, 2), (1, 0), (1, 18)])
def test_batch_equals_each_problem_alone(batch_inputs, precision, flags):
    """flags 2: no graph replay; 16: per-sweep kernels on every coarse level
    (the level-to-kernel mapping never changes a result: block-visit kernels
    round where the per-sweep kernels store)."""
    scenes, rhs = batch_inputs
    alone = _alone(scenes, rhs, precision, flags)
    bt, batched = _batched(scenes, rhs, precision, flags)
    info = bt.info()
    _say(f"  precision {precision} flags {flags}: merged levels {info['nlevels']}, leaves0 {info['leaves0']}, "
         f"iterations {[r[1] for r in batched]}")
    assert info["nproblems"] == len(scenes) and info["nlevels"] == 3
    res = bt.last_residual()
    for i, ((s1, i1, x1, r1), (s2, i2, x2)) in enumerate(zip(alone, batched)):
        assert bt.ndof(i) == x1.size
        assert s1 == s2 == U.SUCCESS, f"problem {i}"
        assert i1 == i2, f"problem {i}: {i1} vs {i2}"
        assert np.array_equal(x1, x2), f"problem {i}"
        # the residual replacement (BF16 form) replaces the same residuals of the same problems
        assert r1[1] == res[i][1] and (r1[0] == res[i][0] or precision == 0), f"problem {i}: {r1} vs {res[i]}"


def test_batch_iteration_limit(batch_inputs):
    scenes, rhs = batch_inputs
    alone = _alone(scenes, rhs, max_iterations=2)
    bt, batched = _batched(scenes, rhs, max_iterations=2)
    for (s1, i1, x1, _), (s2, i2, x2) in zip(alone, batched):
        assert s1 == s2 and i1 == i2 and np.array_equal(x1, x2)


def test_batch_confines_failure_and_trivial_problems(batch_inputs):
    """problem 3 gets a NaN in its right-hand side (it fails at the iteration
    limit), problem 6 a zero right-hand side (converged before iterating); the
    others are unaffected."""
    import torch
    scenes, rhs = batch_inputs
    bad = [b.clone() for b in rhs]
    bad[3][bad[3].numel() // 2] = float("nan")
    bad[6].zero_()
    bt, batched = _batched(scenes, bad, max_iterations=8)
    clean = _alone(scenes, rhs, max_iterations=8)
    assert batched[3][0] == U.FAILED
    assert batched[6][0] == U.SUCCESS and batched[6][1] == 0 and not batched[6][2].any()
    for i in (0, 1, 2, 4, 5):
        assert batched[i][0] == clean[i][0] and batched[i][1] == clean[i][1]
        assert np.array_equal(batched[i][2], clean[i][2]), f"problem {i}"
    torch.cuda.synchronize()


def test_batch_warm_start_and_repeat(batch_inputs):
    """A second solve from the first's solutions (the serving warm start) and a
    repeat solve on the same handle give the per-problem results again."""
    import torch
    scenes, rhs = batch_inputs
    bt, first = _batched(scenes, rhs)
    x0 = [torch.from_numpy(r[2]).to(rhs[0].device) for r in first]
    xs = [x.clone() for x in x0]
    st, it = bt.solve_device([x.data_ptr() for x in xs], [b.data_ptr() for b in rhs])
    torch.cuda.synchronize()
    for i, (s, b) in enumerate(zip(scenes, rhs)):
        lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
        x = x0[i].clone()
        rc, it1 = lap.solve_pcg_device(x.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        assert st[i] == rc and it[i] == it1
        assert torch.equal(xs[i], x), f"problem {i}"


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_oracle(batch_inputs, precision):
    scenes, rhs = batch_inputs
    _, batched = _batched(scenes, rhs, precision)
    for i, (s, b) in enumerate(zip(scenes, rhs)):
        orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
        bn = b.cpu().numpy()
        st_o, it_o, xo = orc.solve_pcg(bn)
        _, _, x64 = orc.solve_pcg64(bn)
        st, it, x = batched[i]
        _say(f"  problem {i} precision {precision}: status {st} / oracle {st_o}, iterations {it} / {it_o}, "
             f"rel L2 vs FP32 oracle {_rel(x, xo):.2e}, vs FP64 {_rel(x, x64):.2e}")
        assert st == st_o == U.SUCCESS, f"problem {i}"
        assert _rel(x, x64) <= 1e-5, f"problem {i}"
        if precision == 0:
            assert abs(it - it_o) <= 1, f"problem {i}"
            assert _rel(x, xo) <= 1e-5, f"problem {i}"
        else:
            # the BF16 W-cycle form has no reference counterpart: its iteration count is not a parity
            # quantity, but it must not need more iterations than the reference's FP32 rule (its FP64
            # residual replacement removes the FP32 recurrence's stagnation, so it often needs fewer)
            assert it <= it_o + 1, f"problem {i}"


# Scene(3, 0, seed=22): a 36^3 problem (24084 DOF, 2 levels) on which the
# reference's PCG rule fails (||r||_inf grows at iteration 5 after iteration 3)
# and so does FLIP's pure-multigrid fallback (growth after iteration 8, at 14);
# the batch reports both and returns the pure-multigrid result.
FAILING = (3, 0, 22)


@pytest.fixture(scope="module")
def failing_inputs(hip_available):
    import torch
    dev = torch.device("cuda:0")
    scenes, rhs = [], []
    for cfg, n, seed in (FAILING, (3, 40, 3), (0, 32, 0)):
        s = U.Scene(cfg, n, seed=seed)
        lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
        scenes.append(s)
        rhs.append(torch.from_numpy(lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)).to(dev))
    return scenes, rhs


def test_failing_problem_is_the_references(failing_inputs):
    """the oracle's PCG and pure multigrid both fail on it, at iterations 5 and 14"""
    scenes, rhs = failing_inputs
    s = scenes[0]
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    b = rhs[0].cpu().numpy()
    st, it, _ = orc.solve_pcg(b)
    assert (st, it) == (U.FAILED, 5)
    st, it, _ = orc.solve_mg(b)
    assert (st, it) == (U.FAILED, 14)


@pytest.mark.parametrize("max_iterations", [100, 1])
def test_batch_pure_multigrid_fallback(failing_inputs, max_iterations):
    """flags bit 7: the failed problem is zeroed and solved by solvePureMultigrid with the same
    tolerance and iteration limit; the others keep their PCG result. max_iterations = 1 makes
    every PCG fail."""
    import torch
    scenes, rhs = failing_inputs
    bt = U.PoissonBatch([(s.phi, s.weights, 1.0, s.dx) for s in scenes], flags=128,
                        max_iterations=max_iterations)
    xs = [torch.zeros_like(b) for b in rhs]
    st, it = bt.solve_device([x.data_ptr() for x in xs], [b.data_ptr() for b in rhs])
    torch.cuda.synchronize()
    fb = bt.last_fallback()
    for i, (s, b) in enumerate(zip(scenes, rhs)):
        orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
        bn = b.cpu().numpy()
        st_o, it_o, xo = orc.solve_pcg(bn, maxit=max_iterations)
        x = xs[i].cpu().numpy()
        _say(f"  problem {i} max_it {max_iterations}: PCG {st[i]}/{it[i]} (oracle {st_o}/{it_o}), "
             f"fallback {fb[i]}")
        assert st[i] == st_o and abs(it[i] - it_o) <= 1, f"problem {i}"
        if st_o == U.SUCCESS:
            assert fb[i] == (-1, -1)
            assert _rel(x, xo) <= 1e-5
            continue
        mst, mit, xm = orc.solve_mg(bn, maxit=max_iterations)
        assert fb[i][0] == mst and abs(fb[i][1] - mit) <= 1, f"problem {i}: {fb[i]} vs {(mst, mit)}"
        if mst == U.SUCCESS:
            assert _rel(x, xm) <= 1e-5, f"problem {i}: {_rel(x, xm):.2e}"
        else:
            # a failed multigrid iteration ends where the residual stops falling: on the seeded 36^3
            # problem an enclosed liquid pocket leaves a near-singular block, along whose constant mode
            # the two iterates drift apart by rounding; what they must share is the residual they reach
            r_ours = np.abs(orc.apply(0, 1, x, bn)).max()
            r_orc = np.abs(orc.apply(0, 1, xm, bn)).max()
            _say(f"  problem {i}: failed multigrid residual {r_ours:.3e} vs the oracle's {r_orc:.3e}")
            assert r_ours <= 2.0 * r_orc and r_orc <= 2.0 * r_ours, f"problem {i}"
    # the same problems without the flag: FAILED results stay the PCG's
    bt0 = U.PoissonBatch([(s.phi, s.weights, 1.0, s.dx) for s in scenes], max_iterations=max_iterations)
    ys = [torch.zeros_like(b) for b in rhs]
    st0, it0 = bt0.solve_device([y.data_ptr() for y in ys], [b.data_ptr() for b in rhs])
    torch.cuda.synchronize()
    assert list(st0) == list(st) and list(it0) == list(it)
    assert all(f == (-1, -1) for f in bt0.last_fallback())
