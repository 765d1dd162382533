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


def test_viscosity_solve_waits_for_the_callers_stream(hip_available):
    import torch
    dev = torch.device("cuda:0")
    v = U.ViscScene(0, 32, seed=0, mu=1.0)
    lw = U.L_with_level(v.viscosity, v.liquid, v.solid, v.dt, v.rho, v.dx)
    vs = U.simd_viscosity3d(lw, v.velocity)
    b = torch.from_numpy(vs.m_rhs.data.copy()).to(dev)
    x = torch.zeros_like(b)
    it0 = vs.solve_pcg_device(1e-6, x.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    ref = x.cpu().numpy()
    x.fill_(1e3)
    torch.cuda.synchronize()
    _busy(torch, dev)
    x.zero_()
    it = vs.solve_pcg_device(1e-6, x.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    _say(f"  viscosity: iterations {it} (reference {it0})")
    assert it == it0 and np.array_equal(x.cpu().numpy(), ref)


def test_device_rhs_waits_for_the_callers_stream(hip_available):
    """grid values already in device memory, written on the null stream just before the call"""
    import torch
    dev = torch.device("cuda:0")
    s = U.Scene(1, 48, seed=1)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
    grids = [s.weights, s.vx, s.vy, s.vz]
    ref = lap.createPressurePoissonRightHandSide(*grids)
    src = [torch.from_numpy(np.ascontiguousarray(g.values, dtype=np.float32)).to(dev) for g in grids]
    dst = [torch.full_like(t, 1e3) for t in src]
    out = torch.full((lap.mNumDof,), float("nan"), device=dev)
    torch.cuda.synchronize()
    _busy(torch, dev)
    for d, t in zip(dst, src):
        d.copy_(t)   # queued behind the products
    lap.createPressurePoissonRightHandSide_device(out.data_ptr(), *grids, values_ptrs=[t.data_ptr() for t in dst])
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)


def test_pressure_gradient_waits_for_the_callers_stream(hip_available):
    """the pressure is written on the null stream just before the update reads it"""
    import torch
    dev = torch.device("cuda:0")
    s = U.Scene(0, 32, seed=0)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
    b = torch.from_numpy(lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)).to(dev)
    p = torch.zeros_like(b)
    lap.solve_pcg_device(p.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    vel = [s.vx, s.vy, s.vz]
    ref = lap.apply_pressure_gradient_device(p.data_ptr(), s.phi, s.weights, None, vel, dt=0.04, friction=0.0)
    q = torch.full_like(p, 1e3)
    torch.cuda.synchronize()
    _busy(torch, dev)
    q.copy_(p)
    got = lap.apply_pressure_gradient_device(q.data_ptr(), s.phi, s.weights, None, vel, dt=0.04, friction=0.0)
    for c in range(3):
        for a, r in zip(got[c], ref[c]):
            assert np.array_equal(a, r), f"channel {c}"


def test_server_waits_for_the_callers_stream(hip_available):
    """a warm start x0 written on the null stream just before submit"""
    import torch
    from uaamg_amd.server import SolveServer
    dev = torch.device("cuda:0")
    s = U.Scene(1, 48, seed=1)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
    b = torch.from_numpy(lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)).to(dev)
    with SolveServer(device=0, slots=2) as srv:
        pid = srv.add_poisson(s.phi, s.weights, 1.0, s.dx)
        x0 = torch.zeros_like(b)
        out0 = torch.zeros_like(b)
        st0, it0, _, _ = srv.wait(srv.submit(pid, b.data_ptr(), out0.data_ptr(), x0.data_ptr()))
        ref = out0.cpu().numpy()
        x0.fill_(1e3)
        out = torch.zeros_like(b)
        torch.cuda.synchronize()
        _busy(torch, dev)
        x0.zero_()
        st, it, _, _ = srv.wait(srv.submit(pid, b.data_ptr(), out.data_ptr(), x0.data_ptr()))
        torch.cuda.synchronize()
        _say(f"  server: status {st}, iterations {it} (reference {it0})")
        assert st == st0 == U.SUCCESS and it == it0
        assert np.array_equal(out.cpu().numpy(), ref)
