"""Stream ordering at the device entry points: a handle on its own stream
waits for the work already queued on the null stream (torch's default
stream) before it reads the vectors handed in. Each case queues a long
matrix product ahead of the kernel that writes the solve's initial guess, so
without that wait the solve would read the guess before it is written."""
import numpy as np
import pytest

import uaamg_amd as U

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _busy(torch, dev):
    a = torch.randn(4096, 4096, device=dev)
    for _ in range(4):
        a = a @ a
        a = a / a.abs().max()
    return a


def test_solve_waits_for_the_callers_stream(hip_available):
    import torch
    dev = torch.device("cuda:0")
    s = U.Scene(1, 48, seed=1)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
    b = torch.from_numpy(lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)).to(dev)
    x = torch.zeros_like(b)
    rc0, it0 = lap.solve_pcg_device(x.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    ref = x.cpu().numpy()
    for k in range(3):
        x.fill_(1e3)
        torch.cuda.synchronize()
        _busy(torch, dev)
        x.zero_()   # queued behind the products on the null stream
        rc, it = lap.solve_pcg_device(x.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        _say(f"  solve {k}: status {rc}, iterations {it} (reference {it0})")
        assert rc == rc0 == U.SUCCESS and it == it0
        assert np.array_equal(x.cpu().numpy(), ref)


def test_batch_waits_for_the_callers_stream(hip_available):
    import torch
    dev = torch.device("cuda:0")
    scenes = [U.Scene(3, 24, seed=1), U.Scene(1, 32, seed=2)]
    bt = U.PoissonBatch([(s.phi, s.weights, 1.0, s.dx) for s in scenes])
    rhs = []
    for s in scenes:
        lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
        rhs.append(torch.from_numpy(lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)).to(dev))
    xs = [torch.zeros_like(b) for b in rhs]
    st0, it0 = bt.solve_device([x.data_ptr() for x in xs], [b.data_ptr() for b in rhs])
    torch.cuda.synchronize()
    ref = [x.cpu().numpy() for x in xs]
    for x in xs:
        x.fill_(1e3)
    torch.cuda.synchronize()
    _busy(torch, dev)
    for x in xs:
        x.zero_()
    st, it = bt.solve_device([x.data_ptr() for x in xs], [b.data_ptr() for b in rhs])
    torch.cuda.synchronize()
    assert st == st0 and it == it0
    for x, r in zip(xs, ref):
        assert np.array_equal(x.cpu().numpy(), r)
