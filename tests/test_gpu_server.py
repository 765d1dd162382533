"""GPU checks of the C++ solve server: every result agrees with the oracle
(Poisson: solveMultigridPCG's status, iterations within 1 and 1e-5 relative L2
against the FP32 oracle and its FP64 twin; viscosity: pcg_solve's iterations
within 1 and 1e-5 against both); each result is the same alone and in the mix,
under a memory budget that forces evictions too, also with four workers
evicting and re-acquiring problems all the time; warm starts come from the
resident (or evicted and restored) solution; the server runs the same solver
as a direct handle (bit-identical to uaamg_poisson_solve_pcg_device); a
problem registered with FLIP's fallback (flags bit 7) is solved again by pure
multigrid when its PCG fails, as the oracle's solvePureMultigrid.
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

pytestmark = pytest.mark.gpu

# (kind, config, n, seed, precision)
PROBLEMS = [("poisson", 3, 40, 3, 0), ("poisson", 1, 48, 1, 1), ("poisson", 3, 64, 7, 0),
            ("viscosity", 3, 32, 5, 0)]


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pool(hip_available):
    import torch
    dev = torch.device("cuda:0")
    out = []
    for kind, cfg, n, seed, prec in PROBLEMS:
        _say(f"[setup] server problem {kind} cfg={cfg} n={n} seed={seed} precision={prec}")
        if kind == "poisson":
            sc = U.Scene(cfg, n, seed=seed)
            lap = U.LaplacianWithLevel(sc.phi, sc.weights, dt=1.0, dx=sc.dx, device=-1)
            b = lap.createPressurePoissonRightHandSide(sc.weights, sc.vx, sc.vy, sc.vz)
            tol = 1e-7
        else:
            sc = U.ViscScene(cfg, n, seed=seed, mu=1.0)
            lw = U.L_with_level(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx, device=-1)
            b = lw.build_rhs(sc.velocity).data
            tol = 1e-6
        g = torch.Generator(device=dev).manual_seed(seed)
        bd = torch.from_numpy(b).to(dev)
        rhs = [bd, bd * (1 + 0.05 * torch.randn(bd.shape, device=dev, generator=g))]
        out.append({"kind": kind, "scene": sc, "prec": prec, "rhs": rhs, "tol": tol})
    return out


def _register(srv, p):
    sc = p["scene"]
    if p["kind"] == "poisson":
        return srv.add_poisson(sc.phi, sc.weights, 1.0, sc.dx, precision=p["prec"])
    return srv.add_viscosity(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx)


def _solve_all(srv, pool, pids, mix: bool, reps: int = 2):
    """Each problem solves rhs[0] then rhs[1] from zero; `mix` submits everything
    before waiting, otherwise one request at a time."""
    import torch
    zeros = [torch.zeros_like(p["rhs"][0]) for p in pool]
    outs = {(i, r, k): torch.zeros_like(p["rhs"][0]) for i, p in enumerate(pool) for r in range(2)
            for k in range(reps)}
    order = [(i, r, k) for k in range(reps) for r in range(2) for i in range(len(pool))]
    res = {}
    pending = []
    for key in order:
        i, r, k = key
        p = pool[i]
        t = srv.submit(pids[i], p["rhs"][r].data_ptr(), outs[key].data_ptr(), zeros[i].data_ptr(), p["tol"])
        if mix:
            pending.append((key, t))
        else:
            res[key] = srv.wait(t)
    for key, t in pending:
        res[key] = srv.wait(t)
    torch.cuda.synchronize()
    return {key: (res[key][0], res[key][1], outs[key].cpu().numpy()) for key in order}


def test_alone_and_in_the_mix_bit_identical(pool):
    from uaamg_amd.server import SolveServer
    with SolveServer(slots=1) as srv:
        pids = [_register(srv, p) for p in pool]
        alone = _solve_all(srv, pool, pids, mix=False)
    with SolveServer(slots=4) as srv:
        pids = [_register(srv, p) for p in pool]
        mixed = _solve_all(srv, pool, pids, mix=True)
        st = srv.stats()
    _say(f"  mix: peak in flight {st['peak_inflight']}, completed {st['completed']}")
    assert st["completed"] == len(mixed) and st["failed"] == 0
    for key, (s, it, x) in alone.items():
        s2, it2, x2 = mixed[key]
        assert s == s2 == 0 and it == it2, key
        assert np.array_equal(x, x2), key


def test_eviction_under_a_tight_budget(pool):
    from uaamg_amd.server import SolveServer
    with SolveServer(slots=1) as srv:
        pids = [_register(srv, p) for p in pool]
        sizes = [srv.problem_info(pid)["device_bytes"] for pid in pids]
        alone = _solve_all(srv, pool, pids, mix=False, reps=1)
    budget = max(sizes) + min(sizes) // 2   # at most one large problem resident at a time
    with SolveServer(slots=3, memory_budget=budget) as srv:
        pids = [_register(srv, p) for p in pool]
        tight = _solve_all(srv, pool, pids, mix=True, reps=1)
        st = srv.stats()
    _say(f"  budget {budget} B: evictions {st['evictions']}, acquisitions {st['acquisitions']}")
    assert st["evictions"] > 0 and st["failed"] == 0
    assert st["resident_bytes"] <= budget
    for key, (s, it, x) in alone.items():
        s2, it2, x2 = tight[key]
        assert s == s2 == 0 and it == it2 and np.array_equal(x, x2), key


def test_eviction_churn_on_four_slots(pool):
    """Four workers and a budget of about two problems: problems are evicted and made resident
    again all the time, so graph captures (a re-acquired problem's first solve) run while other
    workers free memory, copy an evicted solution out or order their streams after the null
    stream. No request may end in an error, and each result is the one the problem gives alone."""
    from uaamg_amd.server import SolveServer
    with SolveServer(slots=1) as srv:
        pids = [_register(srv, p) for p in pool]
        sizes = sorted(srv.problem_info(pid)["device_bytes"] for pid in pids)
        alone = _solve_all(srv, pool, pids, mix=False, reps=1)
    budget = sizes[-1] + sizes[-2]
    with SolveServer(slots=4, memory_budget=budget) as srv:
        pids = [_register(srv, p) for p in pool]
        churn = _solve_all(srv, pool, pids, mix=True, reps=6)
        st = srv.stats()
    _say(f"  budget {budget} B: evictions {st['evictions']}, acquisitions {st['acquisitions']}, "
         f"peak in flight {st['peak_inflight']}")
    errors = {key: r[0] for key, r in churn.items() if r[0] != 0}
    assert not errors and st["failed"] == 0, errors
    assert st["evictions"] > 0
    for (i, r, k), (s, it, x) in churn.items():
        s1, it1, x1 = alone[(i, r, 0)]
        assert s == s1 and it == it1 and np.array_equal(x, x1), (i, r, k)


def test_warm_start_survives_eviction(pool):
    """Solve rhs[0] cold, then rhs[1] warm from the kept solution: the same with
    and without an eviction in between. The tolerance is relative to each
    solve's own initial residual, so the warm solve takes about as many
    iterations as the cold one."""
    import torch
    from uaamg_amd.server import SolveServer
    p0, p1 = pool[0], pool[2]
    results = []
    for evict in (False, True):
        sizes = None
        with SolveServer(slots=1) as probe:
            sizes = [probe.problem_info(_register(probe, p))["device_bytes"] for p in (p0, p1)]
        budget = max(sizes) + 1024 if evict else 0
        with SolveServer(slots=1, memory_budget=budget) as srv:
            a = _register(srv, p0)
            b = _register(srv, p1)
            x = torch.zeros_like(p0["rhs"][0])
            s, it_cold, _, _ = srv.wait(srv.submit(a, p0["rhs"][0].data_ptr(), 0, 0, p0["tol"]))
            assert s == 0
            # the other problem displaces `a` when the budget holds only one
            s, _, _, _ = srv.wait(srv.submit(b, p1["rhs"][0].data_ptr(), 0, 0, p1["tol"]))
            assert s == 0
            assert srv.problem_info(a)["resident"] == (not evict)
            s, it_warm, _, _ = srv.wait(srv.submit(a, p0["rhs"][1].data_ptr(), x.data_ptr(), 0, p0["tol"]))
            assert s == 0
            torch.cuda.synchronize()
            results.append((it_cold, it_warm, x.cpu().numpy()))
    (c0, w0, x0), (c1, w1, x1) = results
    assert c0 == c1 and w0 == w1 and np.array_equal(x0, x1)
    assert w0 <= c0 + 1


def test_server_matches_direct_handle(pool):
    import torch
    from uaamg_amd.server import SolveServer
    p = pool[2]
    sc = p["scene"]
    lap = U.LaplacianWithLevel(sc.phi, sc.weights, dt=1.0, dx=sc.dx)
    x = torch.zeros_like(p["rhs"][0])
    rc, it = lap.solve_pcg_device(x.data_ptr(), p["rhs"][0].data_ptr())
    with SolveServer(slots=2) as srv:
        pid = _register(srv, p)
        y = torch.zeros_like(x)
        s, it2, _, _ = srv.wait(srv.submit(pid, p["rhs"][0].data_ptr(), y.data_ptr(), 0, 1e-7))
    torch.cuda.synchronize()
    assert rc == s == 0 and it == it2
    assert torch.equal(x, y)


def test_server_errors(pool):
    import torch
    from uaamg_amd.server import SolveServer
    with SolveServer(slots=1, memory_budget=1 << 20) as srv:
        pid = _register(srv, pool[0])
        b = pool[0]["rhs"][0]
        with pytest.raises(U.UaamgError):   # footprint above the budget
            srv.submit(pid, b.data_ptr(), 0, 0, 1e-7)
        with pytest.raises(U.UaamgError):   # unknown problem
            srv.submit(pid + 100, b.data_ptr(), 0, 0, 1e-7)
        with pytest.raises(U.UaamgError):
            srv.wait(12345)
    torch.cuda.synchronize()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_server_results_match_oracle(pool):
    import torch
    from uaamg_amd.server import SolveServer
    with SolveServer(slots=4) as srv:
        pids = [_register(srv, p) for p in pool]
        outs = [[torch.zeros_like(p["rhs"][r]) for r in range(2)] for p in pool]
        # x0 = none: request 0 starts from zero, request 1 from request 0's resident solution
        tickets = [[srv.submit(pids[i], p["rhs"][r].data_ptr(), outs[i][r].data_ptr(), 0, p["tol"])
                    for r in range(2)] for i, p in enumerate(pool)]
        res = [[srv.wait(t) for t in ts] for ts in tickets]
    torch.cuda.synchronize()
    for i, p in enumerate(pool):
        sc = p["scene"]
        for r in range(2):
            b = p["rhs"][r].cpu().numpy()
            st, it = res[i][r][0], res[i][r][1]
            x = outs[i][r].cpu().numpy()
            # request r = 1 is warm-started from request 0's solution (the problem's resident one)
            x0 = None if r == 0 else outs[i][0].cpu().numpy()
            if p["kind"] == "poisson":
                orc = O.OraclePoisson(sc.phi, sc.weights, 1.0, sc.dx)
                st_o, it_o, xo = orc.solve_pcg(b, x0=x0)
                _, _, x64 = orc.solve_pcg64(b)
                assert st == st_o == U.SUCCESS
            else:
                orc = O.OracleViscosity(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx)
                it_o, xo, _ = orc.solve(b, x0=x0, tol=p["tol"])
                _, x64 = orc.solve64(b, tol=1e-12, maxit=200)
                assert st == U.SUCCESS
            _say(f"  {p['kind']} {i} request {r}: iterations {it} / oracle {it_o}, rel L2 vs FP32 oracle "
                 f"{_rel(x, xo):.2e}, vs FP64 {_rel(x, x64):.2e}")
            assert _rel(x, x64) <= 1e-5, (i, r)
            if p["prec"] == 0:
                assert abs(it - it_o) <= 1, (i, r)
                assert _rel(x, xo) <= 1e-5, (i, r)
            else:
                # BF16 W-cycle form: no reference counterpart; no more iterations than the FP32 rule
                assert it <= it_o + 1, (i, r)


def test_server_pure_multigrid_fallback(hip_available):
    """Scene(3, 0, seed=22): the reference's PCG fails at iteration 5 and its pure multigrid at 14;
    registered with flags bit 7 the server returns the PCG's FAILED status and iterations, the
    fallback's status and iterations, and the pure-multigrid solution."""
    import torch
    from uaamg_amd.server import SolveServer
    dev = torch.device("cuda:0")
    sc = U.Scene(3, 0, seed=22)
    lap = U.LaplacianWithLevel(sc.phi, sc.weights, dt=1.0, dx=sc.dx, device=-1)
    bn = lap.createPressurePoissonRightHandSide(sc.weights, sc.vx, sc.vy, sc.vz)
    b = torch.from_numpy(bn).to(dev)
    orc = O.OraclePoisson(sc.phi, sc.weights, 1.0, sc.dx)
    st_o, it_o, _ = orc.solve_pcg(bn)
    mst, mit, xm = orc.solve_mg(bn)
    with SolveServer(slots=2) as srv:
        pid = srv.add_poisson(sc.phi, sc.weights, 1.0, sc.dx, flags=128)
        plain = srv.add_poisson(sc.phi, sc.weights, 1.0, sc.dx)
        out = torch.zeros_like(b)
        out0 = torch.zeros_like(b)
        st, it, fst, fit, _, _ = srv.wait_fallback(srv.submit(pid, b.data_ptr(), out.data_ptr(), 0))
        st0, it0, fst0, fit0, _, _ = srv.wait_fallback(srv.submit(plain, b.data_ptr(), out0.data_ptr(), 0))
    torch.cuda.synchronize()
    _say(f"  server fallback: PCG {st}/{it} (oracle {st_o}/{it_o}), MG {fst}/{fit} (oracle {mst}/{mit})")
    assert st == st_o == U.FAILED and abs(it - it_o) <= 1
    assert fst == mst and abs(fit - mit) <= 1
    # both multigrid iterations fail (an enclosed liquid pocket leaves a near-singular block, along
    # whose constant mode the iterates drift apart by rounding): they must reach the same residual
    x = out.cpu().numpy()
    r_ours = np.abs(orc.apply(0, 1, x, bn)).max()
    r_orc = np.abs(orc.apply(0, 1, xm, bn)).max()
    _say(f"  failed multigrid residual {r_ours:.3e} vs the oracle's {r_orc:.3e}")
    assert r_ours <= 2.0 * r_orc and r_orc <= 2.0 * r_ours
    assert (st0, fst0, fit0) == (U.FAILED, -1, -1) and it0 == it
