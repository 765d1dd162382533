"""GPU parity of the merged viscosity batch (uaamg_viscosity_batch_*): every
problem's result and m_iteration are bit-identical to its pcg_solve alone on
its own handle (simd_viscosity3d.cpp:3694-3740), for a batch of uneven
problems (one- to four-level hierarchies, their own dt, rho, dx and
viscosity, given out of depth order), with and without graph replay, at the
iteration limit, from a warm start and on a repeat solve; a zero right-hand
side stays zero and confined to its problem; every result agrees with the
oracle's FP32 pcg_solve (iterations within 1, 1e-5 relative L2 to its FP64
solve).
"""
import numpy as np
import pytest

import oracle as O
from uaamg_amd.viscosity import L_with_level, ViscBatch, ViscScene, simd_viscosity3d

pytestmark = pytest.mark.gpu

# (config, n, seed, mu, dt, rho): hierarchies of 2, 1, 3, 3, 4, 1, 3 levels
PROBLEMS = [(3, 16, 6, 1.0, 1 / 24, 1000.0), (0, 12, 3, 1.0, 1 / 24, 1000.0), (0, 48, 0, 1.0, 1 / 24, 1000.0),
            (3, 40, 5, 1.0, 1 / 30, 800.0), (0, 64, 1, 20.0, 1 / 24, 1000.0), (3, 8, 1, 1.0, 1 / 48, 1000.0),
            (1, 32, 3, 5.0, 1 / 24, 1200.0)]
TOL = 1e-6


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def visc_inputs(hip_available):
    import torch
    dev = torch.device("cuda:0")
    scenes, rhs = [], []
    for cfg, n, seed, mu, dt, rho in PROBLEMS:
        _say(f"[setup] viscosity batch problem cfg={cfg} n={n} seed={seed} mu={mu} dt={dt:.4f} rho={rho}")
        v = ViscScene(cfg, n, seed=seed, mu=mu, dt=dt, rho=rho)
        lw = L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx, device=-1)
        scenes.append(v)
        rhs.append(torch.from_numpy(lw.build_rhs(v.velocity).data).to(dev))
    return scenes, rhs


def _alone(scenes, rhs, flags=0, max_iterations=100, x0=None):
    import torch
    out = []
    for i, (v, b) in enumerate(zip(scenes, rhs)):
        lw = L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx, flags=flags)
        vs = simd_viscosity3d(lw, v.velocity)
        vs.m_max_iter = max_iterations
        x = torch.zeros_like(b) if x0 is None else x0[i].clone()
        it = vs.solve_pcg_device(TOL, x.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        out.append((it, x.cpu().numpy(), lw.nlevels))
    return out


def _batch(scenes, flags=0, max_iterations=100):
    return ViscBatch([(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx) for v in scenes], flags=flags,
                     max_iterations=max_iterations)


def _solve(bt, rhs, x0=None):
    import torch
    xs = [torch.zeros_like(b) if x0 is None else x0[i].clone() for i, b in enumerate(rhs)]
    it = bt.solve_device(TOL, [x.data_ptr() for x in xs], [b.data_ptr() for b in rhs])
    torch.cuda.synchronize()
    return [(it[i], xs[i].cpu().numpy()) for i in range(len(rhs))]


@pytest.mark.parametrize("flags", [0, 2])
def test_visc_batch_equals_each_problem_alone(visc_inputs, flags):
    """flags 2: no graph replay"""
    scenes, rhs = visc_inputs
    alone = _alone(scenes, rhs, flags)
    bt = _batch(scenes, flags)
    batched = _solve(bt, rhs)
    info = bt.info()
    _say(f"  flags {flags}: merged levels {info['nlevels']}, leaves0 {info['leaves0']}, "
         f"levels {[a[2] for a in alone]}, iterations {[r[0] for r in batched]}")
    assert info["nproblems"] == len(scenes) and info["nlevels"] == max(a[2] for a in alone) == 4
    assert sorted({a[2] for a in alone}) == [1, 2, 3, 4]
    for i, ((i1, x1, _), (i2, x2)) in enumerate(zip(alone, batched)):
        assert bt.ndof(i) == x1.size
        assert i1 == i2, f"problem {i}: {i1} vs {i2}"
        assert np.array_equal(x1, x2), f"problem {i}"


@pytest.mark.parametrize("max_iterations", [0, 1, 3])
def test_visc_batch_iteration_limit(visc_inputs, max_iterations):
    """every problem stops at it == m_max_iter - 1 (the reference's rule has no failure status);
    m_max_iter = 0 leaves x as it was"""
    scenes, rhs = visc_inputs
    alone = _alone(scenes, rhs, max_iterations=max_iterations)
    batched = _solve(_batch(scenes, max_iterations=max_iterations), rhs)
    for i, ((i1, x1, _), (i2, x2)) in enumerate(zip(alone, batched)):
        assert i1 == i2, f"problem {i}"
        assert np.array_equal(x1, x2), f"problem {i}"


def test_visc_batch_warm_start_repeat_and_zero_rhs(visc_inputs):
    """A second solve from the first's solutions (the serving warm start) and a repeat solve on
    the same handle give the per-problem results again; problem 1's zero right-hand side leaves
    its zero iterate untouched and changes nothing elsewhere."""
    import torch
    scenes, rhs = visc_inputs
    bt = _batch(scenes)
    first = _solve(bt, rhs)
    x0 = [torch.from_numpy(r[1]).to(rhs[0].device) for r in first]
    warm = _solve(bt, rhs, x0)
    alone = _alone(scenes, rhs, x0=x0)
    for i, ((i1, x1, _), (i2, x2)) in enumerate(zip(alone, warm)):
        assert i1 == i2 and np.array_equal(x1, x2), f"problem {i}"
    again = _solve(bt, rhs)
    for i, ((i1, x1), (i2, x2)) in enumerate(zip(first, again)):
        assert i1 == i2 and np.array_equal(x1, x2), f"problem {i}"
    zr = [b.clone() for b in rhs]
    zr[1].zero_()
    zb = _solve(bt, zr)
    za = _alone(scenes[1:2], zr[1:2])
    assert zb[1][0] == za[0][0] and not zb[1][1].any()
    for i in (0, 2, 3, 4, 5, 6):
        assert zb[i][0] == first[i][0] and np.array_equal(zb[i][1], first[i][1]), f"problem {i}"


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_visc_batch_matches_oracle(visc_inputs):
    scenes, rhs = visc_inputs
    batched = _solve(_batch(scenes), rhs)
    for i in (0, 1, 3, 5):
        v, b = scenes[i], rhs[i].cpu().numpy()
        ov = O.OracleViscosity(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx)
        it_o, xo, _ = ov.solve(b, tol=TOL)
        _, x64 = ov.solve64(b, tol=1e-12, maxit=200)
        it, x = batched[i]
        _say(f"  problem {i}: iterations {it} / oracle {it_o}, rel L2 vs FP32 oracle {_rel(x, xo):.2e}, "
             f"vs FP64 {_rel(x, x64):.2e}")
        assert abs(it - it_o) <= 1, f"problem {i}"
        assert _rel(x, xo) <= 1e-5 and _rel(x, x64) <= 1e-5, f"problem {i}"
