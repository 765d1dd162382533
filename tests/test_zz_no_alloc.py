"""The allocation-free right-hand-side build (uaamg_poisson_build_rhs_device_noalloc)
called through the C ABI: two calls in a row on one GPU, the second on a
smaller problem, each bit-identical to uaamg_poisson_build_rhs_device on the
same device-resident grids and to the host build.
"""
import ctypes as C

import numpy as np
import pytest

import uaamg_amd as U

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _bind():
    L = U.lib()   # declares both entry points' C signatures
    return L.uaamg_poisson_build_rhs_device_noalloc, L.uaamg_poisson_build_rhs_device


def _device_grids(scene, dev):
    """the five input grids with their values copied to the device (solid velocity: none)"""
    import torch
    keep, structs = [], []
    for g in (scene.weights, scene.vx, scene.vy, scene.vz):
        v = torch.from_numpy(np.ascontiguousarray(g.values, np.float32)).to(dev)
        keep.append(v)
        structs.append(g._struct(v.data_ptr()))
    return keep, structs


@pytest.mark.parametrize("sizes", [((1, 64, 1), (0, 32, 0))])
def test_noalloc_rhs_matches_the_allocating_build(hip_available, sizes):
    import torch
    dev = torch.device("cuda:0")
    noalloc, alloc = _bind()
    for cfg, n, seed in sizes:
        _say(f"[case] noalloc rhs build cfg={cfg} n={n} seed={seed}")
        s = U.Scene(cfg, n, seed=seed)
        lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
        keep, st = _device_grids(s, dev)
        ref = torch.full((lap.mNumDof,), float("nan"), device=dev)
        out = torch.full((lap.mNumDof,), float("nan"), device=dev)
        torch.cuda.synchronize()
        args = [C.byref(x) for x in st] + [None]
        rc = alloc(lap._h, *args, 1, C.c_void_p(ref.data_ptr()))
        assert rc == 0, U.lib().uaamg_last_error()
        rc = noalloc(lap._h, *args, C.c_void_p(out.data_ptr()))
        assert rc == 0, U.lib().uaamg_last_error()
        torch.cuda.synchronize()
        _say(f"  ndof {lap.mNumDof}: bit-identical {torch.equal(ref, out)}")
        assert torch.equal(ref, out)
        # and both equal the host build
        host = lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)
        assert np.array_equal(out.cpu().numpy(), host)
        del keep
