"""The mixed-precision viscosity solve (precision 1: the W-cycle's level
vectors below level 0 stored in BF16, FP32 arithmetic, FP32 PCG with FP64 dot
products). It has no reference counterpart, so it is pinned by what the FP32
form already establishes:

- its result agrees with the oracle's FP64 pcg_solve (1e-5 relative L2) and
  its FP64 true residual on the explicit level-0 matrix meets the tolerance
  as closely as the FP32 form's does; it takes no more iterations than the
  oracle's FP32 solve plus one;
- one W-cycle (z = M r) stays within BF16 rounding of the FP32 W-cycle;
- the level-to-kernel mappings (single-workgroup or per-pass half-visits)
  and graph replay give bit-identical results, as in the FP32 form.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from uaamg_amd.viscosity import L_with_level, ViscScene, simd_viscosity3d

pytestmark = pytest.mark.gpu

CASES = [(0, 48, 0, 1.0), (3, 40, 5, 1.0), (0, 64, 1, 20.0)]


def _say(*a):
    print(*a, flush=True)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _solve(v, precision, flags=0, tol=1e-6):
    import torch
    lw = L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx, precision=precision, flags=flags)
    vs = simd_viscosity3d(lw, v.velocity)
    b = torch.from_numpy(vs.m_rhs.data).to("cuda:0")
    x = torch.zeros_like(b)
    it = vs.solve_pcg_device(tol, x.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    return lw, vs, it, x.cpu().numpy()


@pytest.mark.parametrize("cfg,n,seed,mu", CASES)
def test_mixed_viscosity_against_fp64_oracle(hip_available, cfg, n, seed, mu):
    _say(f"[case] viscosity mixed cfg={cfg} n={n} seed={seed} mu={mu}")
    v = ViscScene(cfg, n, seed=seed, mu=mu)
    lw32, vs32, it32, x32 = _solve(v, 0)
    lwbf, vsbf, itbf, xbf = _solve(v, 1)
    ov = O.OracleViscosity(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx)
    b = vs32.m_rhs.data
    it_o, xo, _ = ov.solve(b, tol=1e-6)
    _, x64 = ov.solve64(b, tol=1e-12, maxit=200)
    ptr, col, val = lw32.level_matrix(0)
    A = sp.csr_matrix((val.astype(np.float64), col, ptr), shape=(b.size, b.size))
    b64 = b.astype(np.float64)
    tr32 = np.abs(b64 - A @ x32.astype(np.float64)).max() / np.abs(b64).max()
    trbf = np.abs(b64 - A @ xbf.astype(np.float64)).max() / np.abs(b64).max()
    _say(f"  ndof {b.size}, levels {lwbf.nlevels}: iterations FP32 {it32} / mixed {itbf} / oracle {it_o}; "
         f"rel L2 vs FP64: FP32 {_rel(x32, x64):.2e}, mixed {_rel(xbf, x64):.2e}; FP64 true rel residual "
         f"FP32 {tr32:.2e}, mixed {trbf:.2e}")
    assert abs(it32 - it_o) <= 1
    assert itbf <= it_o + 1
    assert _rel(xbf, x64) <= 1e-5
    assert trbf <= max(2.0 * tr32, 2e-6)


def test_mixed_viscosity_wcycle_within_bf16_rounding(hip_available):
    v = ViscScene(0, 48, seed=0, mu=1.0)
    lw32 = L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx)
    lwbf = L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx, precision=1)
    r = simd_viscosity3d(lw32, v.velocity).m_rhs.data
    z32 = simd_viscosity3d(lw32, v.velocity).precondition(r)
    zbf = simd_viscosity3d(lwbf, v.velocity).precondition(r)
    d = _rel(zbf, z32)
    _say(f"  one W-cycle: rel L2 mixed vs FP32 {d:.2e}")
    assert 0.0 < d <= 2e-2
    # a linear preconditioner up to rounding: M(2r) = 2 M(r) holds exactly in both forms
    zbf2 = simd_viscosity3d(lwbf, v.velocity).precondition(2.0 * r)
    assert np.array_equal(zbf2, 2.0 * zbf)


@pytest.mark.parametrize("flags", [2, 32, 128, 34])
def test_mixed_viscosity_mappings_bit_identical(hip_available, flags):
    """flags 2: no graph replay; 32: single-workgroup half-visits on every level with at most 16
    leaves per channel; 128: per-pass launches everywhere"""
    v = ViscScene(3, 40, seed=5, mu=1.0)
    _, _, it_ref, x_ref = _solve(v, 1)
    _, _, it, x = _solve(v, 1, flags=flags)
    assert it == it_ref and np.array_equal(x, x_ref)
