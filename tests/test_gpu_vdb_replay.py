"""A scene read from an OpenVDB file, solved on the GPU: the reference's
liquid source level set (tests/golden/sphere_source.vdb, read with its interior
tiles expanded) as the liquid SDF of a pressure-Poisson problem with open
faces (an empty face-weight grid of background 1) and seeded face velocities
on its leaves. The right-hand side equals the oracle's bit for bit, the FP32
solve matches the oracle's solveMultigridPCG (status, iterations within 1,
1e-5 relative L2 against it and its FP64 twin), and the BF16 W-cycle form
meets 1e-5 against the FP64 solve.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import uaamg_amd as U
from uaamg_amd.vdb import read_vdb

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "sphere_source.vdb"
DT = 1.0 / 60.0


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def test_sphere_source_poisson_replay(hip_available):
    import torch
    print("[case] vdb replay: sphere_source.vdb as liquid phi, open faces", flush=True)
    g = read_vdb(str(FIXTURE), expand_tiles=128)[0]
    phi, dx = g.grid, g.voxel_size[0]
    w = U.Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512, 3), np.float32), None, (1.0, 1.0, 1.0))
    rng = np.random.default_rng(11)
    vel = [U.Grid(phi.origins.copy(), rng.uniform(-1, 1, (phi.origins.shape[0], 512)).astype(np.float32), None, 0.0)
           for _ in range(3)]
    orc = O.OraclePoisson(phi, w, DT, dx)
    b_o = orc.build_rhs(w, *vel)
    st_o, it_o, xo = orc.solve_pcg(b_o)
    _, _, x64 = orc.solve_pcg64(b_o)
    out = {}
    for prec in (U.PRECISION_FP32, U.PRECISION_BF16_CYCLE):
        lap = U.LaplacianWithLevel(phi, w, dt=DT, dx=dx, precision=prec)
        b = lap.createPressurePoissonRightHandSide(w, *vel)
        assert np.array_equal(b, b_o)
        bd = torch.from_numpy(b).to("cuda:0")
        x = torch.zeros_like(bd)
        rc, it = lap.solve_pcg_device(x.data_ptr(), bd.data_ptr())
        torch.cuda.synchronize()
        out[prec] = (rc, it, x.cpu().numpy())
        print(f"  precision {prec}: ndof {lap.mNumDof}, levels {lap.nlevels}, status {rc}, iterations {it} "
              f"(oracle {st_o}, {it_o}), rel L2 vs FP64 {_rel(out[prec][2], x64):.2e}", flush=True)
    rc, it, x = out[U.PRECISION_FP32]
    assert rc == st_o == U.SUCCESS and abs(it - it_o) <= 1
    assert _rel(x, xo) <= 1e-5 and _rel(x, x64) <= 1e-5
    rc, it, x = out[U.PRECISION_BF16_CYCLE]
    assert rc == U.SUCCESS and _rel(x, x64) <= 1e-5
