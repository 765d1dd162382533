"""GPU parity of the face weights from the solid SDF (uaamg_face_weights_ex: the
static FLIP_vdb.cpp:1990-2068 and the member variant FLIP_vdb.cpp:2669-2752 with its
1e-2 cut): topology, active masks and every weight bit-identical to the oracle's
restatement, for a scene's liquid SDF topology and a sampled solid
sphere SDF (cut faces at every orientation), and for a solid grid missing leaves
(background lookups).
"""
import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

pytestmark = pytest.mark.gpu


def _sphere_sdf(origins: np.ndarray, c, r) -> U.Grid:
    vals = np.zeros((origins.shape[0], 512), np.float32)
    off = np.arange(512)
    lx, ly, lz = off >> 6, (off >> 3) & 7, off & 7
    for i, o in enumerate(origins):
        x, y, z = o[0] + lx, o[1] + ly, o[2] + lz
        vals[i] = (np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r).astype(np.float32)
    return U.Grid(origins.copy(), vals, None, (3.0,))


@pytest.mark.parametrize("cfg,n,seed", [(0, 32, 0), (1, 48, 1)])
def test_face_weights_equal_oracle(hip_available, cfg, n, seed):
    s = U.Scene(cfg, n, seed=seed)
    solid = _sphere_sdf(s.weights.origins, (n * 0.5 + 0.3, n * 0.45, n * 0.55 - 0.2), n * 0.2)
    # drop a tenth of the solid leaves: those lookups read the background
    keep = np.random.default_rng(seed).uniform(size=solid.origins.shape[0]) > 0.1
    thinned = U.Grid(solid.origins[keep].copy(), solid.values[keep].copy(), None, (3.0,))
    for sg in (solid, thinned):
        for member in (False, True):
            got = U.calculate_face_weights(s.phi, sg, member=member)
            org, act, val = O.face_weights(s.phi, sg, cut=member)
            print(f"[case] cfg={cfg} n={n} member={member} leaves {org.shape[0]}", flush=True)
            assert np.array_equal(got.origins, org)
            assert np.array_equal(got.active, act)
            assert np.array_equal(got.values, val)
            cut = ((val > 0) & (val < 1)).sum()
            assert cut > 0   # the sphere cuts faces
            if member:
                assert not ((val > 0) & (val < np.float32(1e-2))).any()
