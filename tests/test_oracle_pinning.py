"""Pins the CPU oracle against independent brute force (no GPU).

Each check is chosen so that a plausible slip in the oracle fails it: a wrong
face weight index or sign changes the dense level-0 matrix; a dropped 0.5 or a
wrong child offset breaks A_c = 0.5 R A P (the reference's own explicit-matrix
identity, CG_benchmarks/main.cpp:1070-1074); a wrong Jacobi/SOR/SPAI0 formula
fails the textbook sweep on the dense matrix; an unsymmetric smoothing
schedule breaks <Mx,y> = <x,My>; a wrong tree order permutes DOF numbers.
"""
from dataclasses import dataclass

import numpy as np
import pytest
import scipy.sparse.linalg as spl

import oracle as O
from tests import _brute as B

DT = 0.37
DX = 0.125


@dataclass
class G:
    origins: np.ndarray
    values: np.ndarray
    active: np.ndarray | None
    background: tuple


def make(seed, nleaf=6, max_coarsest=30, negative=True, weights=(0.0, 0.35, 1.0)):
    rng = np.random.default_rng(seed)
    org, phi, act, w = B.random_problem(rng, nleaf=nleaf, negative=negative, weight_levels=weights)
    phi_bg, w_bg = 0.3, (1.0, 1.0, 1.0)
    orc = O.OraclePoisson(G(org, phi, act, (phi_bg,)), G(org, w, None, w_bg), DT, DX, max_coarsest=max_coarsest)
    A, coords = B.assemble_level0(org, phi, act, w, phi_bg, w_bg, DT, DX)
    return orc, A, coords


SEEDS = [1, 2, 3, 4]


def level_matrix(orc, level):
    n = orc.ndof(level)
    return B.dense_from_apply(lambda e: orc.apply(level, 0, e), n)


def test_fraction_inside_closed_forms():
    f2, f4 = O.fraction_inside2, O.fraction_inside4
    assert f2(-1, -2) == 1 and f2(1, 2) == 0
    assert f2(-1, 1) == pytest.approx(0.5) and f2(-1, 3) == pytest.approx(0.25) and f2(3, -1) == pytest.approx(0.25)
    assert f4(-1, -1, -1, -1) == 1 and f4(1, 1, 1, 1) == 0
    assert f4(-1, -1, 1, 1) == pytest.approx(0.5)          # bottom edge inside, horizontal cut
    assert f4(-1, 1, 1, 1) == pytest.approx(0.125)         # one corner: triangle 1/2 * 1/2 * 1/2
    assert f4(1, -1, -1, -1) == pytest.approx(0.875)


def _exact_halfplane_area(a, b, c):
    """Area of {(u,v) in [0,1]^2 : a + b u + c v < 0} by polygon clipping."""
    poly = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
    f = lambda p: a + b * p[0] + c * p[1]
    out = []
    for i in range(4):
        p, q = poly[i], poly[(i + 1) % 4]
        fp, fq = f(p), f(q)
        if fp < 0:
            out.append(p)
        if (fp < 0) != (fq < 0):
            s = fp / (fp - fq)
            out.append((p[0] + s * (q[0] - p[0]), p[1] + s * (q[1] - p[1])))
    if len(out) < 3:
        return 0.0
    x = np.array([p[0] for p in out])
    y = np.array([p[1] for p in out])
    return 0.5 * abs(np.dot(x, np.roll(y, 1)) - np.dot(y, np.roll(x, 1)))


def test_fraction_inside_linear_fields_exact():
    # for a linear SDF the edge-interpolation formulas give the exact cut area
    rng = np.random.default_rng(0)
    for _ in range(400):
        a, b, c = rng.uniform(-1, 1, 3)
        bl, br, tl, tr = a, a + b, a + c, a + b + c
        got = O.fraction_inside4(bl, br, tl, tr)
        assert got == pytest.approx(_exact_halfplane_area(a, b, c), abs=2e-6)


@pytest.mark.parametrize("seed", SEEDS)
def test_level0_matrix_matches_dense_assembly(seed):
    orc, A, coords = make(seed)
    assert orc.ndof(0) == len(coords)
    M = level_matrix(orc, 0)
    assert np.allclose(M, A, rtol=2e-6, atol=1e-6 * np.abs(A).max())
    assert np.allclose(M, M.T, rtol=1e-6, atol=1e-9)
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0


@pytest.mark.parametrize("seed", SEEDS)
def test_dof_numbering_is_tree_order(seed):
    orc, A, coords = make(seed)
    org, mask, base = orc.topology(0)
    bits = np.unpackbits(mask.view(np.uint8), axis=1, bitorder="little")
    got = []
    for i in range(org.shape[0]):
        for off in np.nonzero(bits[i])[0]:
            lx, ly, lz = B.local(int(off))
            got.append((int(org[i, 0]) + lx, int(org[i, 1]) + ly, int(org[i, 2]) + lz))
    assert got == coords
    counts = bits.sum(1)
    assert np.array_equal(base, np.concatenate([[0], np.cumsum(counts)[:-1]]))


@pytest.mark.parametrize("seed", SEEDS)
def test_coarse_topology_and_galerkin_identity(seed):
    orc, A, coords = make(seed)
    assert orc.nlevels >= 2
    fine_coords = coords
    Af = level_matrix(orc, 0)
    for level in range(1, orc.nlevels):
        coarse, parent = B.coarsen_coords(fine_coords)
        assert orc.ndof(level) == len(coarse)
        org, mask, base = orc.topology(level)
        bits = np.unpackbits(mask.view(np.uint8), axis=1, bitorder="little")
        got = []
        for i in range(org.shape[0]):
            for off in np.nonzero(bits[i])[0]:
                lx, ly, lz = B.local(int(off))
                got.append((int(org[i, 0]) + lx, int(org[i, 1]) + ly, int(org[i, 2]) + lz))
        assert got == coarse
        P = B.prolongation_matrix(parent, len(coarse))
        Ac_expected = 0.5 * (P.T / 8.0) @ Af @ P
        Ac = level_matrix(orc, level)
        scale = np.abs(Ac_expected).max()
        assert np.allclose(Ac, Ac_expected, rtol=3e-5, atol=3e-6 * scale)
        fine_coords, Af = coarse, Ac


def test_interior_coarse_stencil_closed_form():
    # a solid block of liquid: interior coarse voxels carry 6 t_c / -t_c with t_c = t/4
    org = np.array([[x, y, z] for x in (0, 8, 16) for y in (0, 8, 16) for z in (0, 8, 16)], np.int32)
    n = org.shape[0]
    phi = -np.ones((n, 512), np.float32)
    w = np.ones((n, 512, 3), np.float32)
    orc = O.OraclePoisson(G(org, phi, None, (0.5,)), G(org, w, None, (1.0, 1.0, 1.0)), DT, DX, max_coarsest=100)
    assert orc.nlevels >= 2
    t1 = DT / (2 * DX) ** 2
    diag, kept = orc.coefficients(1, 0)
    xs, _ = orc.coefficients(1, 1)
    org1, mask1, _ = orc.topology(1)
    centre = [i for i in range(org1.shape[0]) if tuple(org1[i]) == (0, 0, 0)][0]
    off = (3 << 6) | (3 << 3) | 3
    assert diag[centre, off] == pytest.approx(6 * t1, rel=1e-6)
    assert xs[centre, off] == pytest.approx(-t1, rel=1e-6)


@pytest.mark.parametrize("seed", SEEDS[:2])
def test_smoothers_match_textbook(seed):
    orc, A, coords = make(seed)
    rng = np.random.default_rng(seed + 100)
    for level in range(orc.nlevels):
        M = level_matrix(orc, level)
        n = M.shape[0]
        x = rng.uniform(-1, 1, n).astype(np.float32)
        b = rng.uniform(-1, 1, n).astype(np.float32)
        d = np.diag(M)
        inv = np.where(d != 0, 1 / np.where(d != 0, d, 1), 0)
        scale = np.abs(M).max() * 4
        for w in (2.2473, 6 / 7):
            jac = x + w * inv * (b - M @ x)
            assert np.allclose(orc.apply(level, 2, x, b, w), jac, rtol=1e-5, atol=1e-5 * scale * np.abs(inv).max())
        assert np.allclose(orc.apply(level, 1, x, b), b - M @ x, rtol=1e-5, atol=1e-6 * scale)
        # SPAI0: a_kk / ||a_k||^2 (init_SPAI0)
        s = d / (M * M).sum(1)
        assert np.allclose(orc.apply(level, 3, x, b), x + s * (b - M @ x), rtol=1e-4, atol=1e-5)
        # red-black SOR colour passes on global (x+y+z) parity
        org, mask, _ = orc.topology(level)
        bits = np.unpackbits(mask.view(np.uint8), axis=1, bitorder="little")
        par = []
        for i in range(org.shape[0]):
            for off in np.nonzero(bits[i])[0]:
                lx, ly, lz = B.local(int(off))
                par.append((int(org[i].sum()) + lx + ly + lz) % 2)
        red = np.array(par) == 0
        for omega, mode, sel in ((1.2, 4, red), (1.2, 5, ~red), (1.0, 4, red)):
            off_sum = M @ x - d * x
            upd = (1 - omega) * x + omega * inv * (b - off_sum)
            exp = np.where(sel, upd, x)
            assert np.allclose(orc.apply(level, mode, x, b, omega), exp, rtol=1e-5, atol=1e-5)
        # first sweeps from a zero guess
        assert np.allclose(orc.first_sweep(level, 0, b, 2.2473), 2.2473 * inv * b, rtol=1e-6, atol=1e-9)
        assert np.allclose(orc.first_sweep(level, 1, b, 0), s * b, rtol=1e-4, atol=1e-9)
        y_red = np.where(red, 1.2 * inv * b, 0.0)
        y_black = np.where(~red, -0.2 * y_red + 1.2 * inv * (b - (M @ y_red - d * y_red)), y_red)
        assert np.allclose(orc.first_sweep(level, 2, b, 1.2), y_black, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("seed", SEEDS[:2])
def test_transfers_match_dense(seed):
    orc, A, coords = make(seed)
    rng = np.random.default_rng(seed + 7)
    fine_coords = coords
    for level in range(orc.nlevels - 1):
        coarse, parent = B.coarsen_coords(fine_coords)
        P = B.prolongation_matrix(parent, len(coarse))
        f = rng.uniform(-1, 1, len(fine_coords)).astype(np.float32)
        c = rng.uniform(-1, 1, len(coarse)).astype(np.float32)
        assert np.allclose(orc.restriction(level, f), (P.T @ f) / 8, rtol=1e-6, atol=1e-7)
        for alpha in (1.0, 0.5):
            assert np.allclose(orc.prolongation(level, c, f, alpha), f + alpha * (P @ c), rtol=1e-6, atol=1e-7)
        fine_coords = coarse


@pytest.mark.parametrize("seed", SEEDS)
def test_coarsest_matches_textbook_pcg(seed):
    orc, A, coords = make(seed, max_coarsest=200)
    L = orc.nlevels - 1
    Mc = level_matrix(orc, L)
    b = np.random.default_rng(seed).uniform(-1, 1, Mc.shape[0]).astype(np.float32)
    got = orc.coarsest_solve(b)
    exp = B.eigen_cg(Mc, b)
    assert np.linalg.norm(got - exp) <= 2e-4 * np.linalg.norm(exp)


def _dense_wcycle(levels, P, r, smoother_w=(2.2473, 0.8571, 0.5296), level=0):
    """Dense restatement of muCyclePreconditioner<2,true> with n=3 scheduled Jacobi."""
    A = levels[level]
    if level == len(levels) - 1:
        return B.eigen_cg(A, r)
    d = np.diag(A)
    inv = np.where(d != 0, 1 / np.where(d != 0, d, 1), 0)
    x = smoother_w[0] * inv * r
    for w in smoother_w[1:]:
        x = x + w * inv * (r - A @ x)
    rc = (P[level].T @ (r - A @ x)) / 8
    xc = _dense_wcycle(levels, P, rc, smoother_w, level + 1)
    if level + 1 < len(levels) - 1:
        Ac = levels[level + 1]
        dc = np.diag(Ac)
        invc = np.where(dc != 0, 1 / np.where(dc != 0, dc, 1), 0)
        # second visit: a full cycle started from xc (pre-smoothing not skipped)
        xc = _cycle_from(levels, P, rc, xc, smoother_w, level + 1)
    x = x + P[level] @ xc
    for w in smoother_w[::-1]:
        x = x + w * inv * (r - A @ x)
    return x


def _cycle_from(levels, P, r, x, smoother_w, level):
    A = levels[level]
    if level == len(levels) - 1:
        return B.eigen_cg(A, r)
    d = np.diag(A)
    inv = np.where(d != 0, 1 / np.where(d != 0, d, 1), 0)
    for w in smoother_w:
        x = x + w * inv * (r - A @ x)
    rc = (P[level].T @ (r - A @ x)) / 8
    xc = _dense_wcycle(levels, P, rc, smoother_w, level + 1)
    if level + 1 < len(levels) - 1:
        xc = _cycle_from(levels, P, rc, xc, smoother_w, level + 1)
    x = x + P[level] @ xc
    for w in smoother_w[::-1]:
        x = x + w * inv * (r - A @ x)
    return x


@pytest.mark.parametrize("seed", SEEDS[:3])
def test_wcycle_matches_dense_restatement(seed):
    orc, A, coords = make(seed, nleaf=8, max_coarsest=40)
    levels, P = [], []
    fine_coords = coords
    for level in range(orc.nlevels):
        levels.append(level_matrix(orc, level))
        if level + 1 < orc.nlevels:
            coarse, parent = B.coarsen_coords(fine_coords)
            P.append(B.prolongation_matrix(parent, len(coarse)))
            fine_coords = coarse
    r = np.random.default_rng(seed).uniform(-1, 1, levels[0].shape[0]).astype(np.float32)
    got = orc.precondition(r, 0)
    exp = _dense_wcycle(levels, P, r.astype(np.float64))
    assert np.linalg.norm(got - exp) <= 1e-4 * np.linalg.norm(exp)


@pytest.mark.parametrize("seed", SEEDS[:2])
def test_preconditioner_symmetric_when_coarsest_is_exact(seed):
    # coarsest <= 10 DOF: 10 CG iterations solve it exactly, so M is a fixed
    # linear operator; the mirrored smoothing schedule makes it symmetric
    orc, A, coords = make(seed, nleaf=5, max_coarsest=8)
    assert orc.ndof(orc.nlevels - 1) <= 10
    n = orc.ndof(0)
    rng = np.random.default_rng(seed)
    for sm in (0, 1, 2, 3):
        x = rng.uniform(-1, 1, n).astype(np.float32)
        y = rng.uniform(-1, 1, n).astype(np.float32)
        mx, my = orc.precondition(x, sm), orc.precondition(y, sm)
        lhs, rhs = float(mx @ y), float(x @ my)
        assert abs(lhs - rhs) <= 2e-4 * max(abs(lhs), abs(rhs)), f"smoother {sm}"
        assert float(mx @ x) > 0


@pytest.mark.parametrize("seed", SEEDS)
def test_pcg_converges_to_sparse_direct_solution(seed):
    orc, A, coords = make(seed, nleaf=8, max_coarsest=60)
    b = np.random.default_rng(seed).uniform(-1, 1, A.shape[0]).astype(np.float32)
    xs = np.linalg.solve(A, b.astype(np.float64))
    for sm in (0, 1, 2, 3):
        st, it, x = orc.solve_pcg(b, smoother=sm)
        assert st == 0
        assert np.linalg.norm(x - xs) <= 1e-5 * np.linalg.norm(xs)
    # the FP64 twin solves the same float-coefficient matrix in double precision
    M = level_matrix(orc, 0)
    xm = np.linalg.solve(M, b.astype(np.float64))
    st, it, x64 = orc.solve_pcg64(b.astype(np.float64), tol=1e-13)
    assert np.linalg.norm(x64 - xm) <= 1e-9 * np.linalg.norm(xm)


def test_pcg_stops_on_relative_infinity_norm():
    orc, A, coords = make(9, nleaf=8, max_coarsest=60)
    b = np.random.default_rng(0).uniform(-1, 1, A.shape[0]).astype(np.float32)
    st, it, x, hist = orc.solve_pcg(b, tol=1e-3, history=True)
    assert st == 0
    assert hist[-1] <= 1e-3 * hist[0] < hist[-2]


def test_max_iteration_and_nan_fail():
    orc, A, coords = make(5, nleaf=8, max_coarsest=60)
    b = np.random.default_rng(0).uniform(-1, 1, A.shape[0]).astype(np.float32)
    st, it, _ = orc.solve_pcg(b, maxit=1)
    assert st == 1 and it == 1
    bb = b.copy()
    bb[3] = np.nan
    st, it, _ = orc.solve_pcg(bb, maxit=6)
    assert st == 1


def test_scene_system_matches_scipy():
    import uaamg_amd as U
    s = U.Scene(1, 32, seed=2)
    orc = O.OraclePoisson(s.phi, s.weights, 1.0, s.dx)
    b = orc.build_rhs(s.weights, s.vx, s.vy, s.vz)
    st, it, x = orc.solve_pcg(b)
    assert st == 0
    st64, it64, x64 = orc.solve_pcg64(b.astype(np.float64))
    r = b - orc.apply64(x64)
    assert np.abs(r).max() <= 1e-9 * np.abs(b).max()
    assert np.linalg.norm(x - x64) <= 1e-5 * np.linalg.norm(x64)


def _grid_lookup(grid):
    idx = {tuple(int(v) for v in o): i for i, o in enumerate(grid.origins)}
    bg = grid.background if isinstance(grid.background, (tuple, list)) else (grid.background,) * 3

    def get(g, comp=None):
        o = tuple(int(v) & ~7 for v in g)
        i = idx.get(o)
        off = ((g[0] & 7) << 6) | ((g[1] & 7) << 3) | (g[2] & 7)
        if i is None:
            return np.float32(bg[0 if comp is None else comp])
        v = grid.values[i, off]
        return np.float32(v if comp is None else v[comp])
    return get


@pytest.mark.parametrize("friction", [0.0, 0.5])
def test_pressure_gradient_closed_forms(friction):
    """The oracle's apply_pressure_gradient restatement (FLIP_vdb.cpp:2571-2667) on a
    linear pressure p = 2x + 3y - 5z (integer, exact in FP32) against the update rule
    evaluated face by face from coordinates: interior liquid faces (theta = 1) get
    du = -dt * a_c / dx exactly; faces to an air voxel use theta = phi_below /
    (phi_below - phi_this) clamped at 0.02 with the air pressure 0; weight-0 faces
    and faces with no DOF on either side are turned off and keep their value;
    partially solid faces blend towards the solid velocity."""
    import uaamg_amd as U
    sc = U.Scene(1, 32, seed=1)   # the solid sphere gives fractional face weights
    lap = U.LaplacianWithLevel(sc.phi, sc.weights, dt=1.0, dx=sc.dx, device=-1)
    orc = O.OraclePoisson(sc.phi, sc.weights, 1.0, sc.dx)
    org, mask, base = lap.level_topology(0)
    dof = {}
    pres = np.zeros(lap.mNumDof, np.float32)
    for i, o in enumerate(org):
        k = int(base[i])
        for off in range(512):
            if (int(mask[i, off >> 6]) >> (off & 63)) & 1:
                g = (int(o[0]) + (off >> 6), int(o[1]) + ((off >> 3) & 7), int(o[2]) + (off & 7))
                dof[g] = k
                pres[k] = 2 * g[0] + 3 * g[1] - 5 * g[2]
                k += 1
    rng = np.random.default_rng(3)
    solid = U.Grid(sc.weights.origins.copy(), rng.uniform(-1, 1, sc.weights.values.shape).astype(np.float32), None,
                   (0.0, 0.0, 0.0))
    vel = [sc.vx, sc.vy, sc.vz]
    ones = [np.full((g.origins.shape[0], 8), np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64) for g in vel]
    dt, dx = np.float32(0.5), np.float32(sc.dx)
    out = orc.apply_pressure_gradient(pres, sc.phi, sc.weights, solid,
                                      [(g.origins, a, g.values) for g, a in zip(vel, ones)], float(dt), friction)
    phi, w, sv = _grid_lookup(sc.phi), _grid_lookup(sc.weights), _grid_lookup(solid)
    slope = (2, 3, -5)
    counts = {"interior": 0, "air": 0, "off": 0, "blend": 0}
    for c in range(3):
        act, val, upd, refl = out[c]
        for i in range(0, vel[c].origins.shape[0], 3):   # a third of the leaves keeps the test short
            o = vel[c].origins[i]
            for off in range(0, 512, 7):
                g = (int(o[0]) + (off >> 6), int(o[1]) + ((off >> 3) & 7), int(o[2]) + (off & 7))
                lo = list(g)
                lo[c] -= 1
                lo = tuple(lo)
                u0 = np.float32(vel[c].values[i, off])
                on = (int(act[i, off >> 6]) >> (off & 63)) & 1
                wt = w(g, c)
                if wt <= 0 or (g not in dof and lo not in dof):
                    assert not on and upd[i, off] == u0 and val[i, off] == u0
                    counts["off"] += 1
                    continue
                assert on
                pt = np.float32(pres[dof[g]]) if g in dof else np.float32(0)
                pb = np.float32(pres[dof[lo]]) if lo in dof else np.float32(0)
                ft, fb = phi(g), phi(lo)
                theta = np.float32(1)
                if ft >= 0 or fb >= 0:
                    theta = np.float32(O.fraction_inside2(float(fb), float(ft)))
                    theta = max(theta, np.float32(0.02))
                    counts["air"] += 1
                else:
                    assert pt - pb == slope[c]
                    counts["interior"] += 1
                du = np.float32(np.float32(np.float32(-dt * (pt - pb)) / dx) / theta)
                assert upd[i, off] == du
                up = np.float32(u0 + du)
                assert refl[i, off] == np.float32(up + du)
                if wt < 1:
                    s = np.float32(np.float32(1 - wt) * np.float32(friction))
                    up = np.float32(np.float32(np.float32(1 - s) * up) + np.float32(s * sv(g, c)))
                    counts["blend"] += 1
                assert val[i, off] == up
    assert min(counts.values()) > 0, counts


def _random_topology(rng, nleaf=6, span=3):
    """leaves at random (possibly negative) origins with random active masks"""
    seen = set()
    org = []
    while len(org) < nleaf:
        o = tuple(int(v) * 8 for v in rng.integers(-span, span, 3))
        if o not in seen:
            seen.add(o)
            org.append(o)
    org = np.array(org, np.int32)
    act = np.zeros((nleaf, 8), np.uint64)
    for i in range(nleaf):
        bits = rng.uniform(size=512) < 0.03
        for off in np.nonzero(bits)[0]:
            act[i, off >> 6] |= np.uint64(1) << np.uint64(off & 63)
    return org, act


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_face_weight_topology_is_box_dilation(seed):
    """The weight grid's topology (dilateActiveValues(2, NN_FACE_EDGE_VERTEX) of the liquid SDF):
    oracle and library both against a dense scipy binary dilation with a 5x5x5 box."""
    import scipy.ndimage as ndi
    import uaamg_amd as U
    rng = np.random.default_rng(seed)
    org, act = _random_topology(rng)
    liquid = U.Grid(org, np.zeros((org.shape[0], 512), np.float32), act, (0.0,))
    lo = org.min(axis=0) - 16
    dims = tuple(int(v) for v in org.max(axis=0) + 24 - lo)
    vol = np.zeros(dims, bool)
    for i, o in enumerate(org):
        for off in range(512):
            if (int(act[i, off >> 6]) >> (off & 63)) & 1:
                vol[o[0] - lo[0] + (off >> 6), o[1] - lo[1] + ((off >> 3) & 7), o[2] - lo[2] + (off & 7)] = True
    dil = ndi.binary_dilation(vol, structure=np.ones((5, 5, 5), bool))
    want = {tuple(int(v) for v in o) for o in org}
    for x, y, z in zip(*np.nonzero(dil)):
        want.add((int(x + lo[0]) & ~7, int(y + lo[1]) & ~7, int(z + lo[2]) & ~7))
    o_org, o_act, o_val = O.face_weights(liquid, U.Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32),
                                                        None, (1.0,)))
    assert {tuple(int(v) for v in o) for o in o_org} == want
    for i, o in enumerate(o_org):
        for off in range(0, 512, 3):
            g = (o[0] - lo[0] + (off >> 6), o[1] - lo[1] + ((off >> 3) & 7), o[2] - lo[2] + (off & 7))
            inside = all(0 <= g[k] < dims[k] for k in range(3)) and dil[g]
            assert bool((int(o_act[i, off >> 6]) >> (off & 63)) & 1) == inside
    # the library's host topology pass (no device needed without output values)
    n = __import__("ctypes").c_int64()
    C = __import__("ctypes")
    gl, gs = liquid._struct(), U.Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32), None, (1.0,))._struct()
    assert U.lib().uaamg_face_weights(C.byref(gl), C.byref(gs), -1, C.byref(n), None, None, None) == 0
    assert n.value == o_org.shape[0]


def test_face_weights_closed_form_plane():
    """A solid half-space z < h (SDF z - h, h = 3.25): x- and y-faces are cut at height h - z,
    so w = 1 - clamp(h - z, 0, 1); a z-face lies in one plane: w = 0 below h, 1 above;
    inactive voxels hold 1."""
    import uaamg_amd as U
    h = np.float32(3.25)
    org = np.array([[0, 0, 0], [0, 0, 8]], np.int32)
    act = np.zeros((2, 8), np.uint64)
    act[0, 3] = np.uint64(1) << np.uint64(4 * 8 + 2)     # voxel (3, 4, 2) of the first leaf
    liquid = U.Grid(org, np.zeros((2, 512), np.float32), act, (0.0,))
    sorg = np.array([[x, y, z] for x in range(-8, 24, 8) for y in range(-8, 24, 8) for z in range(-8, 24, 8)], np.int32)
    sval = np.zeros((sorg.shape[0], 512), np.float32)
    for i, o in enumerate(sorg):
        for off in range(512):
            sval[i, off] = np.float32(o[2] + (off & 7)) - h
    org_o, act_o, val_o = O.face_weights(liquid, U.Grid(sorg, sval, None, (1.0,)))
    checked = 0
    for i, o in enumerate(org_o):
        for off in range(512):
            z = int(o[2]) + (off & 7)
            w = val_o[i, off]
            if not (int(act_o[i, off >> 6]) >> (off & 63)) & 1:
                assert tuple(w) == (1.0, 1.0, 1.0)
                continue
            cut = np.float32(min(max(h - np.float32(z), np.float32(0)), np.float32(1)))
            assert w[0] == np.float32(1) - cut and w[1] == np.float32(1) - cut
            assert w[2] == (np.float32(0) if z < h else np.float32(1))
            checked += 1
    assert checked == 125   # the 5x5x5 box around the one active voxel


def test_face_weights_member_variant_cut():
    """The member variant (FLIP_vdb.cpp:2669-2752) equals the static one with every weight below 1e-2
    set to 0: a solid half-space z < 3.995 leaves x/y faces at z = 3 a sliver 0 < w < 1e-2 (cut to 0)
    and every other weight as the static variant's."""
    import uaamg_amd as U
    h = np.float32(3.995)
    org = np.array([[0, 0, 0]], np.int32)
    act = np.zeros((1, 8), np.uint64)
    act[0, 3] = np.uint64(1) << np.uint64(4 * 8 + 2)     # voxel (3, 4, 2)
    liquid = U.Grid(org, np.zeros((1, 512), np.float32), act, (0.0,))
    sorg = np.array([[x, y, z] for x in range(-8, 24, 8) for y in range(-8, 24, 8) for z in range(-8, 24, 8)], np.int32)
    sval = np.zeros((sorg.shape[0], 512), np.float32)
    for i, o in enumerate(sorg):
        for off in range(512):
            sval[i, off] = np.float32(o[2] + (off & 7)) - h
    solid = U.Grid(sorg, sval, None, (1.0,))
    o0, a0, v0 = O.face_weights(liquid, solid)
    o1, a1, v1 = O.face_weights(liquid, solid, cut=True)
    assert np.array_equal(o0, o1) and np.array_equal(a0, a1)
    sliver = (v0 > 0) & (v0 < np.float32(1e-2))
    assert sliver.sum() == 2 * 25          # the x and y faces of the 25 active voxels at z = 3
    assert np.array_equal(v1, np.where(sliver, np.float32(0), v0))
