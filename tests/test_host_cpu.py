"""CPU-side checks of the product library (no GPU needed).

- libuaamg_amd.so loads and exports every function include/uaamg_capi.h declares;
- the product's host hierarchy build (run through a device=-1 handle, which
  only inspects) equals the oracle's bit for bit: per-level leaf origins, DOF
  masks, DOF numbering, the four stencil coefficient grids with their trimmed
  leaves, SPAI0, and the pressure right-hand side (with and without a solid
  velocity);
- solves on a device=-1 handle fail loudly (no CPU fallback);
- scene face weights obey their closed forms.
"""
import re
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import uaamg_amd as U

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    text = (ROOT / "include" / "uaamg_capi.h").read_text()
    names = set(re.findall(r"\b(uaamg_[a-z0-9_]+)\s*\(", text))
    assert len(names) > 20
    lib = U.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


CASES = [(0, 32, 0), (1, 48, 1), (1, 64, 3), (3, 0, 7), (3, 0, 11)]


@pytest.mark.parametrize("cfg,n,seed", CASES)
def test_host_hierarchy_bit_exact_vs_oracle(cfg, n, seed):
    s = U.Scene(cfg, n, seed=seed)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    assert lap.nlevels == orc.nlevels
    for l in range(lap.nlevels):
        po, pm, pb = lap.level_topology(l)
        oo, om, ob = orc.topology(l)
        assert np.array_equal(po, oo) and np.array_equal(pm, om) and np.array_equal(pb, ob), f"level {l}"
        assert lap.level_info(l)["ndof"] == orc.ndof(l)
        assert lap.level_info(l)["nleaf_tree"] == orc.nleaf_tree(l)
        for kind in range(5):
            pc, pk = lap.level_coefficients(l, kind)
            oc, ok = orc.coefficients(l, kind)
            assert np.array_equal(pk, ok), f"level {l} kind {kind}: trimmed leaves differ"
            assert np.array_equal(pc, oc), f"level {l} kind {kind}: coefficients differ"
    rp = lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)
    ro = orc.build_rhs(s.weights, s.vx, s.vy, s.vz)
    assert np.array_equal(rp, ro)
    # with a solid velocity (non-zero background, leaves on the weights' topology)
    rng = np.random.default_rng(seed + 100)
    sv = U.Grid(s.weights.origins.copy(),
                rng.uniform(-1, 1, (s.weights.origins.shape[0], 512, 3)).astype(np.float32), None,
                (0.5, -0.25, 0.125))
    rp = lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz, sv)
    ro = orc.build_rhs(s.weights, s.vx, s.vy, s.vz, sv)
    assert np.array_equal(rp, ro)


def test_leaf_order_independent_of_input_order():
    s = U.Scene(0, 32, seed=0)
    perm = np.random.default_rng(0).permutation(s.phi.origins.shape[0])
    phi2 = U.Grid(s.phi.origins[perm], s.phi.values[perm], s.phi.active[perm], s.phi.background)
    w2 = U.Grid(s.weights.origins[perm], s.weights.values[perm], s.weights.active[perm], s.weights.background)
    a = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
    b = U.LaplacianWithLevel(phi2, w2, dt=1.0, dx=s.dx, device=-1)
    for l in range(a.nlevels):
        for x, y in zip(a.level_topology(l), b.level_topology(l)):
            assert np.array_equal(x, y)


def test_host_only_handle_refuses_to_solve():
    s = U.Scene(0, 32, seed=0)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
    b = lap.createPressurePoissonRightHandSide(s.weights, s.vx, s.vy, s.vz)
    solver = U.PoissonSolver(lap)
    with pytest.raises(U.UaamgError):
        solver.solveMultigridPCG(lap.getZeroVectorGrid(), b)
    with pytest.raises(U.UaamgError):
        lap.laplacianApply(np.zeros(lap.mNumDof, np.float32))


def test_bad_inputs_rejected():
    s = U.Scene(0, 32, seed=0)
    bad = U.Grid(s.phi.origins + 1, s.phi.values, s.phi.active, s.phi.background)
    with pytest.raises(U.UaamgError):
        U.LaplacianWithLevel(bad, s.weights, dt=1.0, dx=s.dx, device=-1)
    with pytest.raises(U.UaamgError):
        U.LaplacianWithLevel(s.phi, s.weights, dt=-1.0, dx=s.dx, device=-1)
    with pytest.raises(U.UaamgError):   # a Vec3f grid where a float grid belongs
        U.LaplacianWithLevel(s.weights, s.weights, dt=1.0, dx=s.dx, device=-1)


def test_problems_without_dofs_rejected():
    """No liquid voxel (all phi >= 0, or no leaves at all): both creates report UAAMG_EINVAL
    ("no degrees of freedom") rather than building an empty system."""
    w = U.Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512, 3), np.float32), None, (1.0, 1.0, 1.0))
    dry = U.Grid(np.zeros((1, 3), np.int32), np.ones((1, 512), np.float32), None, 1.0)
    empty = U.Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32), None, 1.0)
    for phi in (dry, empty):
        with pytest.raises(U.UaamgError, match="no degrees of freedom"):
            U.LaplacianWithLevel(phi, w, dt=1.0, dx=0.1, device=-1)
        with pytest.raises(U.UaamgError, match="no degrees of freedom"):
            U.L_with_level(empty, phi, empty, 1.0 / 24.0, 1000.0, 0.1, device=-1)


def test_fraction_inside_product_equals_oracle():
    lib = U.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        v = rng.uniform(-1, 1, 4).astype(np.float32)
        assert lib.uaamg_fraction_inside4(*v) == O.fraction_inside4(*v)
        assert lib.uaamg_fraction_inside2(*v[:2]) == O.fraction_inside2(*v[:2])


def test_scene_face_weights_closed_forms():
    s = U.Scene(0, 32, seed=0)   # no solid: interior faces open, wall faces closed
    n = s.n
    w = s.weights.values
    for i in range(s.phi.origins.shape[0]):
        o = s.phi.origins[i]
        for off in (0, 7, 63, 64 * 7 + 63, 300):
            g = (o[0] + (off >> 6), o[1] + ((off >> 3) & 7), o[2] + (off & 7))
            for c in range(3):
                inside = all(0 <= g[k] < n for k in range(3))
                open_face = inside and g[c] - 1 >= 0
                assert w[i, off, c] == (1.0 if open_face else 0.0)
    s1 = U.Scene(1, 48, seed=1)   # solid sphere: fractional weights in [0,1], some strictly between
    w1 = s1.weights.values
    assert w1.min() >= 0 and w1.max() <= 1
    assert np.any((w1 > 0) & (w1 < 1))


def test_scene_is_deterministic():
    a = U.Scene(3, 0, seed=5)
    b = U.Scene(3, 0, seed=5)
    assert a.n == b.n and 32 <= a.n <= 192
    assert np.array_equal(a.phi.values, b.phi.values) and np.array_equal(a.vx.values, b.vx.values)


@pytest.mark.parametrize("cfg,n,seed", [(1, 48, 1), (3, 0, 7)])
def test_triplets_are_the_level_operator(cfg, n, seed):
    """getTriplets of every level, as a sparse matrix in FP64, applies like the
    oracle's FP64 operator on level 0 and like its FP32 operator on every level
    (to FP32 rounding); it is symmetric, and each row lists its diagonal first."""
    import scipy.sparse as sp
    s = U.Scene(cfg, n, seed=seed)
    lap = U.LaplacianWithLevel(s.phi, s.weights, dt=1.0, dx=s.dx, device=-1)
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    rng = np.random.default_rng(3)
    for l in range(lap.nlevels):
        r, c, v = lap.getTriplets(l)
        nd = lap.level_ndof(l)
        first = np.r_[True, r[1:] != r[:-1]]
        assert np.array_equal(r[first], np.arange(nd)) and np.array_equal(c[first], np.arange(nd))
        A = sp.csr_matrix((v.astype(np.float64), (r, c)), shape=(nd, nd))
        assert abs(A - A.T).max() == 0.0, f"level {l} not symmetric"
        x = rng.uniform(-1, 1, nd).astype(np.float32)
        y = A @ x.astype(np.float64)
        y32 = orc.apply(l, 0, x)
        assert np.allclose(y, y32, rtol=0, atol=1e-5 * np.abs(y).max()), f"level {l}"
        if l == 0:
            assert np.allclose(y, orc.apply64(x.astype(np.float64)), rtol=1e-12, atol=1e-12 * np.abs(y).max())
