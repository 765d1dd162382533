"""GPU parity of the device pressure-gradient update
(uaamg_poisson_apply_pressure_gradient_device, FLIP_vdb.cpp:2571-2667): active
masks, updated velocities, update and reflected grids bit-identical to the
oracle's restatement, for the pressure of an actual solve, with and without a
solid velocity, friction, host- and device-resident values.
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _solid(grid: U.Grid, seed: int) -> U.Grid:
    rng = np.random.default_rng(seed)
    vals = rng.uniform(-1.0, 1.0, size=(grid.origins.shape[0], 512, 3)).astype(np.float32)
    return U.Grid(grid.origins.copy(), vals, None, (0.5, -0.25, 0.125))


@pytest.mark.parametrize("cfg,n,seed", [(0, 32, 0), (1, 48, 1)])
def test_gradient_equals_oracle(hip_available, cfg, n, seed):
    import torch
    dev = torch.device("cuda:0")
    _say(f"[setup] gradient cfg={cfg} n={n} seed={seed}")
    s = U.Scene(cfg, n, seed=seed)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx)
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    b = lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)
    x = torch.zeros(lap.mNumDof, dtype=torch.float32, device=dev)
    lap.solve_pcg_device(x.data_ptr(), torch.from_numpy(b).to(dev).data_ptr())
    torch.cuda.synchronize()
    p = x.cpu().numpy()
    vel = [s.vx, s.vy, s.vz]
    velo = [(g.origins, g.active, g.values) for g in vel]
    for solid, friction in ((None, 0.0), (_solid(s.weights, seed + 5), 0.35)):
        ref = orc.apply_pressure_gradient(p, s.phi, s.weights, solid, velo, dt=0.04, friction=friction)
        for on_dev in (False, True):
            got = lap.apply_pressure_gradient_device(x.data_ptr(), s.phi, s.weights, solid, vel, dt=0.04,
                                                     friction=friction, values_on_device=on_dev)
            for c in range(3):
                ra, rv, ru, rr = ref[c]
                ga, gv, gu, gr = got[c]
                tag = f"solid={solid is not None} device={on_dev} channel {c}"
                assert np.array_equal(ga, ra), tag + ": active"
                assert np.array_equal(gv, rv.reshape(gv.shape)), tag + ": values"
                assert np.array_equal(gu, ru.reshape(gu.shape)), tag + ": update"
                assert np.array_equal(gr, rr.reshape(gr.shape)), tag + ": reflected"
            on = sum(int(np.unpackbits(ga.view(np.uint8)).sum()) for ga, _, _, _ in got)
            _say(f"  solid={solid is not None} device={on_dev}: {on} faces updated")
