"""Pins the viscosity oracle (oracle/viscosity_oracle.cpp) against what the
mathematics fixes, independently of its own formulas.

- volume_fraction of a cube is exact for linear SDFs (the two 5-tetrahedron
  decompositions integrate a piecewise-linear interpolant, which equals the
  linear function): axis planes give a, diagonal planes c^2/2 and c^3/6;
- a planar free surface gives closed-form voxel / face / edge volumes (the
  trilinear subcell sampling of a linear SDF is exact);
- in fully liquid interior the 15-point stencil equals its constant-viscosity
  closed form (1 + 8 f mu on the diagonal, -2 f mu along the channel axis,
  -f mu across it, +-f mu cross terms), f = dt / (rho dx^2);
- the explicit matrix is symmetric and positive definite;
- a uniform velocity field is a null vector of the viscous part: with
  u = u_solid = U everywhere, A U equals the right-hand side (this ties the
  operator, the DOF / solid marking and the solid-boundary terms of build_rhs
  together: a dropped or mis-signed term breaks it);
- the coarse operator is the Galerkin product A_c = R A P with P the
  piecewise-constant injection and R = P^T / 8 (untrimmed finest level);
- restriction and prolongation are adjoint up to the factor 8;
- every applier mode equals its textbook form on the explicit CSR matrix, and
  the XYZ / ZYX 6-colour Gauss-Seidel sweeps equal a CSR multicolour
  Gauss-Seidel in the same colour order;
- the coarsest solve equals a float64 textbook Jacobi-PCG (10 iterations);
- the W-cycle preconditioner is symmetric, and PCG converges to the direct
  solution of the explicit system.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle as O
import uaamg_amd as U


# ------------------------------------------------------------ geometry
def _corners(f):
    # argument order of volume_fraction: phi000, phi100, phi010, phi110, phi001, phi101, phi011, phi111
    return [f(i, j, k) for (i, j, k) in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1),
                                         (0, 1, 1), (1, 1, 1)]]


def test_volume_fraction_extremes():
    assert O.volume_fraction8([-1.0] * 8) == 1.0
    assert O.volume_fraction8([1.0] * 8) == 0.0
    assert O.volume_fraction8([0.5] * 8) == 0.0


@pytest.mark.parametrize("a", [0.1, 0.25, 0.5, 0.7, 0.93])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_volume_fraction_axis_plane(a, axis):
    # liquid where coordinate < a
    v = O.volume_fraction8(_corners(lambda i, j, k: (i, j, k)[axis] - a))
    assert abs(v - a) < 2e-6
    # and where coordinate > a
    v = O.volume_fraction8(_corners(lambda i, j, k: a - (i, j, k)[axis]))
    assert abs(v - (1 - a)) < 2e-6


@pytest.mark.parametrize("c", [0.2, 0.5, 0.9])
def test_volume_fraction_diagonal_planes(c):
    v2 = O.volume_fraction8(_corners(lambda i, j, k: i + j - c))
    assert abs(v2 - c * c / 2) < 2e-6
    v3 = O.volume_fraction8(_corners(lambda i, j, k: i + j + k - c))
    assert abs(v3 - c ** 3 / 6) < 2e-6


def test_volume_fraction_matches_product_restatement():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        p = rng.uniform(-1, 1, 8).astype(np.float32)
        assert O.volume_fraction8(p) == U.lib().uaamg_volume_fraction8(p.ctypes.data)


# ---------------------------------------------------- planar free surface
def _box_grid(lo, hi, fn, background):
    """Leaves covering [lo, hi)^3 (multiples of 8) with values fn(x, y, z), all active."""
    org = []
    for x in range(lo, hi, 8):
        for y in range(lo, hi, 8):
            for z in range(lo, hi, 8):
                org.append((x, y, z))
    org = np.array(org, np.int32)
    off = np.arange(512)
    lx, ly, lz = off >> 6, (off >> 3) & 7, off & 7
    vals = np.stack([fn(o[0] + lx, o[1] + ly, o[2] + lz) for o in org]).astype(np.float32)
    return U.Grid(org, vals, None, (background,) * 3)


@pytest.fixture(scope="module")
def pool():
    n, h = 32, 13.3   # surface at y = 13.3 voxels
    dx = 1.0 / n
    liquid = _box_grid(-8, 40, lambda x, y, z: ((y - h) * dx).astype(np.float32), 3 * dx)
    solid = U.Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32), None, (3 * dx,) * 3)
    mu = U.Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32), None, (1.0,) * 3)
    dt, rho = 1.0 / 24, 1000.0
    orc = O.OracleViscosity(mu, liquid, solid, dt, rho, dx, trim=False)
    return dict(n=n, h=h, dx=dx, orc=orc, liquid=liquid, solid=solid, mu=mu, dt=dt, rho=rho)


def test_planar_volumes_closed_form(pool):
    orc, h = pool["orc"], pool["h"]
    ys = np.arange(8, 20)
    pts = np.array([(12, y, 17) for y in ys], np.int32)
    clip = lambda t: np.clip(t, 0.0, 1.0)
    # voxel centre v spans [y - 1/2, y + 1/2]; face of channel 1 at v spans [y - 1, y]
    np.testing.assert_allclose(orc.volume_at(0, pts), clip(h - (ys - 0.5)), atol=2e-6)
    np.testing.assert_allclose(orc.volume_at(1, pts), clip(h - (ys - 0.5)), atol=2e-6)   # face x: same y span
    np.testing.assert_allclose(orc.volume_at(2, pts), clip(h - (ys - 1.0)), atol=2e-6)   # face y
    np.testing.assert_allclose(orc.volume_at(3, pts), clip(h - (ys - 0.5)), atol=2e-6)   # face z
    # edges: edge c at v lies at v - 1/2 on the two other axes
    np.testing.assert_allclose(orc.volume_at(4, pts), clip(h - (ys - 1.0)), atol=2e-6)   # edge x (y shifted)
    np.testing.assert_allclose(orc.volume_at(5, pts), clip(h - (ys - 0.5)), atol=2e-6)   # edge y (y not shifted)
    np.testing.assert_allclose(orc.volume_at(6, pts), clip(h - (ys - 1.0)), atol=2e-6)   # edge z


def test_interior_stencil_closed_form(pool):
    orc = pool["orc"]
    f = pool["dt"] / (pool["rho"] * pool["dx"] ** 2)
    org, mask, base = orc.topology(0)
    # an interior voxel well below the surface and away from the box edges
    v = np.array([16, 5, 16], np.int32)
    li = np.where((org == (v // 8) * 8).all(axis=1))[0][0]
    off = ((v[0] % 8) << 6) | ((v[1] % 8) << 3) | (v[2] % 8)
    for c in range(3):
        d, _ = orc.coefficients(0, c)
        assert abs(d[li, off] - (1 + 8 * f)) <= 1e-6 * (1 + 8 * f)
        for a in range(3):
            g, _ = orc.coefficients(0, 6 + 3 * c + a)
            want = -2 * f if a == c else -f
            assert abs(g[li, off] - want) <= 1e-6 * abs(want)
    for i in range(3):
        for j in range(2):
            for k in range(2):
                g, _ = orc.coefficients(0, 15 + 4 * i + 2 * j + k)
                want = -f if (j + k) % 2 else f
                assert abs(g[li, off] - want) <= 1e-6 * abs(want)


# -------------------------------------------------------------- scenes
SCENES = [(0, 24, 0, 1.0), (1, 24, 1, 20.0), (3, 16, 5, 50.0)]


@pytest.fixture(scope="module", params=SCENES, ids=[f"cfg{c}-n{n}-mu{m:g}" for c, n, s, m in SCENES])
def scene(request):
    cfg, n, seed, mu = request.param
    s = U.ViscScene(cfg, n, seed=seed, mu=mu)
    orc = O.OracleViscosity(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, trim=False)
    orc_t = O.OracleViscosity(s.viscosity, s.liquid, s.solid, s.dt, s.rho, s.dx, trim=True)
    return s, orc, orc_t


def _csr(orc, level):
    ptr, col, val = orc.matrix(level)
    n = orc.level_ndof(level)
    return sp.csr_matrix((val.astype(np.float64), col, ptr), shape=(n, n))


def _channel_of_dof(orc, level):
    org, mask, base = orc.topology(level)
    ch = np.zeros(orc.level_ndof(level), np.int64)
    for i in range(org.shape[0]):
        for c in range(3):
            k = int(sum(bin(int(w)).count("1") for w in mask[i, c]))
            if k:
                ch[base[i, c]:base[i, c] + k] = c
    return ch


def _dof_coords(orc, level):
    org, mask, base = orc.topology(level)
    xyz = np.zeros((orc.level_ndof(level), 3), np.int64)
    for i in range(org.shape[0]):
        for c in range(3):
            bits = np.unpackbits(mask[i, c].view(np.uint8), bitorder="little")
            offs = np.nonzero(bits)[0]
            if offs.size:
                d = base[i, c] + np.arange(offs.size)
                xyz[d] = org[i] + np.stack([offs >> 6, (offs >> 3) & 7, offs & 7], axis=1)
    return xyz


def test_matrix_symmetric_positive_definite(scene):
    s, orc, orc_t = scene
    for o in (orc, orc_t):
        for l in range(o.nlevels):
            A = _csr(o, l)
            assert abs(A - A.T).max() == 0.0
            # symmetric ordering, no pivoting: the pivots are those of L D L^T,
            # all positive exactly when A is positive definite
            lu = spl.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                          options=dict(SymmetricMode=True))
            assert np.all(lu.U.diagonal() > 0)


def test_uniform_velocity_is_in_the_kernel_of_the_viscous_part(scene):
    s, orc, _ = scene
    U3 = np.array([0.7, -1.3, 0.4], np.float32)
    vel = U.Grid(s.velocity.origins, np.broadcast_to(U3, s.velocity.values.shape).copy(), None, tuple(U3))
    rhs = orc.build_rhs(vel, vel)
    A = _csr(orc, 0)
    uvec = U3[_channel_of_dof(orc, 0)].astype(np.float64)
    Au = A @ uvec
    # Rows along the edges of the container box couple to faces beyond two
    # walls, which the reference neither marks as DOFs nor as solid (they lie
    # outside the once-dilated liquid leaves): their velocity counts as zero
    # there. Every other row, walls included, must satisfy A U = b exactly.
    n = int(round(1.0 / s.dx))
    xyz = _dof_coords(orc, 0)
    nb = ((xyz <= 0) | (xyz >= n - 1)).sum(axis=1)
    keep = nb <= 1
    assert keep.sum() > 0.9 * len(keep)
    assert np.abs(Au - rhs)[keep].max() <= 2e-5 * np.abs(rhs).max()
    # a different solid velocity breaks it (the boundary terms are live)
    zero = U.Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512, 3), np.float32), None, (0.0, 0.0, 0.0))
    rhs0 = orc.build_rhs(vel, zero)
    assert np.abs(Au - rhs0)[keep].max() > 1e-3 * np.abs(rhs).max()


def test_galerkin_coarse_operator(scene):
    s, orc, _ = scene
    for l in range(orc.nlevels - 1):
        A = _csr(orc, l)
        Ac = _csr(orc, l + 1)
        fxyz, cxyz = _dof_coords(orc, l), _dof_coords(orc, l + 1)
        fch, cch = _channel_of_dof(orc, l), _channel_of_dof(orc, l + 1)
        index = {(int(c), *map(int, p)): d for d, (c, p) in enumerate(zip(cch, cxyz))}
        rows = np.arange(len(fch))
        cols = np.array([index[(int(c), *(int(v) >> 1 for v in p))] for c, p in zip(fch, fxyz)])
        P = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(len(fch), len(cch)))
        G = (P.T @ A @ P) / 8.0
        D = (G - Ac).tocoo()
        assert np.abs(D.data).max(initial=0.0) <= 1e-5 * np.abs(Ac.data).max()


def test_transfers_adjoint(scene):
    s, orc, _ = scene
    rng = np.random.default_rng(2)
    for l in range(orc.nlevels - 1):
        f = rng.uniform(-1, 1, orc.level_ndof(l)).astype(np.float32)
        c = rng.uniform(-1, 1, orc.level_ndof(l + 1)).astype(np.float32)
        Rf = orc.restriction(l, f).astype(np.float64)
        Pc = orc.prolongation(l, c, np.zeros_like(f)).astype(np.float64)
        assert abs(Rf @ c - (f.astype(np.float64) @ Pc) / 8.0) <= 1e-5 * (abs(Rf) @ abs(c))


def test_apply_modes_vs_csr(scene):
    s, orc, _ = scene
    rng = np.random.default_rng(4)
    for l in range(orc.nlevels):
        A = _csr(orc, l)
        d = A.diagonal()
        n = A.shape[0]
        x = rng.uniform(-1, 1, n).astype(np.float32)
        b = rng.uniform(-1, 1, n).astype(np.float32)
        x64, b64 = x.astype(np.float64), b.astype(np.float64)
        Ax = A @ x64
        scale = np.abs(A).max() * 2
        assert np.abs(orc.apply(l, 0, x) - Ax).max() <= 1e-5 * scale
        assert np.abs(orc.apply(l, 1, x, b) - (b64 - Ax)).max() <= 1e-5 * scale
        jac = x64 + (b64 - Ax) / d
        assert np.abs(orc.apply(l, 2, x, b) - jac).max() <= 1e-5 * np.abs(jac).max()
        # SPAI0: m_ii = a_ii / sum_j a_ij^2 over the explicit row
        m = d / np.asarray(A.multiply(A).sum(axis=1)).ravel()
        spai = x64 + m * (b64 - Ax)
        assert np.abs(orc.apply(l, 3, x, b) - spai).max() <= 1e-5 * np.abs(spai).max(), f"level {l} SPAI0"
        # 6-colour Gauss-Seidel in the reference's colour order
        xyz, ch = _dof_coords(orc, l), _channel_of_dof(orc, l)
        red = (xyz.sum(axis=1) % 2) == 0
        for mode, order in ((6, [(0, 1), (0, 0), (1, 1), (1, 0), (2, 1), (2, 0)]),
                            (7, [(2, 0), (2, 1), (1, 0), (1, 1), (0, 0), (0, 1)])):
            y = x64.copy()
            for c, is_red in order:
                sel = (ch == c) & (red == bool(is_red))
                r = b64 - A @ y + d * y
                y[sel] = r[sel] / d[sel]
            got = orc.apply(l, mode, x, b)
            assert np.abs(got - y).max() <= 1e-4 * max(np.abs(y).max(), 1.0), f"level {l} mode {mode}"


def test_coarsest_vs_textbook_cg(scene):
    s, orc, orc_t = scene
    o = orc_t
    A = _csr(o, o.nlevels - 1)
    b = np.random.default_rng(6).uniform(-1, 1, A.shape[0])
    inv = 1.0 / A.diagonal()
    x = np.zeros_like(b)
    r = b.copy()
    p = inv * r
    rz = r @ p
    thr = max((np.finfo(np.float32).eps ** 2) * (b @ b), np.finfo(np.float32).tiny)
    for _ in range(10):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if r @ r < thr:
            break
        z = inv * r
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
    got = o.coarsest_solve(b.astype(np.float32))
    assert np.linalg.norm(got - x) <= 1e-5 * np.linalg.norm(x)


def test_preconditioner_symmetric(scene):
    s, orc, orc_t = scene
    rng = np.random.default_rng(8)
    n = orc_t.ndof
    a = rng.uniform(-1, 1, n).astype(np.float32)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    Ma = orc_t.precondition(a).astype(np.float64)
    Mb = orc_t.precondition(b).astype(np.float64)
    lhs, rhs = Ma @ b, a @ Mb
    assert abs(lhs - rhs) <= 1e-4 * (np.abs(Ma) @ np.abs(b))


def test_pcg_converges_to_direct_solution(scene):
    s, orc, orc_t = scene
    b = orc_t.build_rhs(s.velocity, None)
    A = _csr(orc_t, 0)
    xd = spl.spsolve(A.tocsc(), b.astype(np.float64))
    it, x, hist = orc_t.solve(b, tol=1e-7)
    assert hist[-1] <= 1e-7 * hist[0] or it == 99
    assert np.linalg.norm(x - xd) <= 1e-5 * np.linalg.norm(xd)
    it64, x64 = orc_t.solve64(b.astype(np.float64), tol=1e-10)
    assert np.linalg.norm(x64 - xd) <= 1e-8 * np.linalg.norm(xd)


def test_trimmed_operator_close_to_untrimmed(scene):
    s, orc, orc_t = scene
    A, At = _csr(orc, 0), _csr(orc_t, 0)
    assert A.shape == At.shape
    D = abs(A - At)
    assert D.max() <= 1e-3 * abs(A).max() * 1.01
