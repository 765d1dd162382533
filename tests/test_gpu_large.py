"""The Poisson path at its largest configured size (config 1 at 512^3, about
47M DOF and 6 levels), checked through size-independent properties, since the
CPU oracle takes minutes to build at this size: both forms converge; the
FP64 true residual of each result, evaluated by torch's sparse FP64 SpMV on
the explicit level-0 matrix (getTriplets, checked against the oracle's
operators on CPU), meets 1e-7 relative for the BF16 form with residual
replacement and stays near it for the FP32 form; and the two forms' solutions
agree.
"""
import numpy as np
import pytest

import uaamg_amd as U

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_512_cube_both_forms(hip_available):
    import torch
    dev = torch.device("cuda:0")
    _say("[setup] poisson config 1 at 512^3")
    sc = U.Scene(1, 512, seed=1)
    results = {}
    for prec in (U.PRECISION_FP32, U.PRECISION_BF16_CYCLE):
        lap = U.LaplacianWithLevel(sc.phi, sc.weights, dt=1.0, dx=sc.dx, precision=prec)
        if prec == U.PRECISION_FP32:
            b = torch.from_numpy(lap.createPressurePoissonRightHandSide(sc.weights, sc.vx, sc.vy, sc.vz)).to(dev)
            r, c, v = lap.getTriplets(0)
            n = lap.mNumDof
            _say(f"  ndof {n}, levels {lap.nlevels}, triplets {r.size}")
            # triplets come row by row in DOF order, so the row pointer is a prefix count
            assert np.all(r[1:] >= r[:-1])
            crow = np.zeros(n + 1, np.int64)
            np.cumsum(np.bincount(r, minlength=n), out=crow[1:])
            A = torch.sparse_csr_tensor(torch.from_numpy(crow).to(dev), torch.from_numpy(c.astype(np.int64)).to(dev),
                                        torch.from_numpy(v.astype(np.float64)).to(dev), size=(n, n))
            del r, c, v, crow
        x = torch.zeros_like(b)
        rc, it = lap.solve_pcg_device(x.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        res = (b.double() - A @ x.double()).abs().max().item()
        rel = res / b.double().abs().max().item()
        _say(f"  precision {prec}: status {rc}, iterations {it}, true rel residual {rel:.3e}, "
             f"reported {lap.last_residual()}")
        assert rc == U.SUCCESS
        results[prec] = (x, rel, it)
        del lap
    x32, rel32, it32 = results[U.PRECISION_FP32]
    xbf, relbf, itbf = results[U.PRECISION_BF16_CYCLE]
    assert relbf <= 1.01e-7
    assert rel32 <= 1e-6
    assert itbf <= it32 + 4
    diff = (x32.double() - xbf.double()).norm().item() / xbf.double().norm().item()
    _say(f"  FP32 vs BF16 solution: rel L2 {diff:.3e}")
    assert diff <= 1e-5
