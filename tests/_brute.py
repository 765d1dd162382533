"""Independent brute-force restatements for pinning the oracle on tiny problems.

Everything here is assembled voxel by voxel from the definitions (dense
numpy matrices, Python dictionaries), sharing no code with the oracle or the
product.
"""
from __future__ import annotations

import numpy as np


def tree_key(o):
    """OpenVDB Tree4<5,4,3> traversal key of a leaf origin."""
    x, y, z = (int(v) for v in o)
    root = (x & ~4095, y & ~4095, z & ~4095)
    upper = (((x & 4095) >> 7) << 10) | (((y & 4095) >> 7) << 5) | ((z & 4095) >> 7)
    lower = (((x & 127) >> 3) << 8) | (((y & 127) >> 3) << 4) | ((z & 127) >> 3)
    return root + (upper, lower)


def local(off):
    return off >> 6, (off >> 3) & 7, off & 7


def random_problem(rng, nleaf=6, spread=2, negative=True, p_active=0.95, weight_levels=(0.0, 0.35, 1.0)):
    """Random tiny leaf topology with phi, active masks and face weights."""
    lo = -spread if negative else 0
    cells = set()
    while len(cells) < nleaf:
        cells.add(tuple(int(v) for v in rng.integers(lo, spread + 1, 3)))
    origins = np.array(sorted(cells), dtype=np.int32) * 8
    n = origins.shape[0]
    phi = rng.uniform(-1.0, 0.6, (n, 512)).astype(np.float32)
    act_bits = rng.random((n, 512)) < p_active
    active = np.packbits(act_bits, axis=1, bitorder="little").view(np.uint64).reshape(n, 8)
    w = rng.choice(np.array(weight_levels, dtype=np.float32), size=(n, 512, 3), p=[0.1, 0.15, 0.75])
    return origins, phi, active, w


def voxel_maps(origins, values, active=None):
    m = {}
    n = origins.shape[0]
    for i in range(n):
        o = origins[i]
        bits = None
        if active is not None:
            bits = np.unpackbits(active[i].view(np.uint8), bitorder="little")
        for off in range(512):
            lx, ly, lz = local(off)
            c = (int(o[0]) + lx, int(o[1]) + ly, int(o[2]) + lz)
            m[c] = (values[i, off], True if bits is None else bool(bits[off]))
    return m


def fraction_inside(a, b):
    if a < 0 and b < 0:
        return 1.0
    if a < 0 <= b:
        return a / (a - b)
    if a >= 0 > b:
        return b / (b - a)
    return 0.0


def assemble_level0(origins, phi, active, w, phi_bg, w_bg, dt, dx):
    """Dense level-0 matrix from the ghost-fluid definitions.

    Returns (A, dof_coords) with DOFs ordered by (tree key of the leaf, offset).
    """
    pm = voxel_maps(origins, phi, active)
    wm = voxel_maps(origins, w)
    t = dt / (dx * dx)

    def getphi(c):
        return float(pm[c][0]) if c in pm else float(phi_bg)

    def getw(c, comp):
        return float(wm[c][0][comp]) if c in wm else float(w_bg[comp])

    diag = {}
    for c, (p, act) in pm.items():
        if not act or not (p < 0):
            continue
        d = 0.0
        for comp in range(3):
            for sgn in (+1, -1):
                nc = list(c)
                nc[comp] += sgn
                nc = tuple(nc)
                wt = getw(nc if sgn > 0 else c, comp)
                other = getphi(nc)
                term = wt * t
                if other < 0:
                    d += term
                else:
                    theta = max(fraction_inside(p, other), 0.02)
                    d += term / theta
        if d != 0.0:
            diag[c] = d
    coords = sorted(diag, key=lambda c: (tree_key((c[0] & ~7, c[1] & ~7, c[2] & ~7)),
                                         ((c[0] & 7) << 6) | ((c[1] & 7) << 3) | (c[2] & 7)))
    index = {c: i for i, c in enumerate(coords)}
    n = len(coords)
    A = np.zeros((n, n))
    for c, i in index.items():
        A[i, i] = diag[c]
        for comp in range(3):
            for sgn in (+1, -1):
                nc = list(c)
                nc[comp] += sgn
                nc = tuple(nc)
                j = index.get(nc)
                if j is None:
                    continue
                wt = getw(nc if sgn > 0 else c, comp)
                A[i, j] = -wt * t
    return A, coords


def order_coords(coords):
    return sorted(coords, key=lambda c: (tree_key((c[0] & ~7, c[1] & ~7, c[2] & ~7)),
                                         ((c[0] & 7) << 6) | ((c[1] & 7) << 3) | (c[2] & 7)))


def coarsen_coords(fine_coords):
    """Coarse DOFs = voxels with at least one fine child, and the fine->coarse map."""
    coarse = order_coords({(c[0] // 2, c[1] // 2, c[2] // 2) for c in fine_coords})
    cidx = {c: i for i, c in enumerate(coarse)}
    parent = np.array([cidx[(c[0] // 2, c[1] // 2, c[2] // 2)] for c in fine_coords])
    return coarse, parent


def prolongation_matrix(parent, ncoarse):
    P = np.zeros((len(parent), ncoarse))
    P[np.arange(len(parent)), parent] = 1.0
    return P


def dense_from_apply(apply_fn, n):
    """Columns of an operator applied to unit vectors."""
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n, np.float32)
        e[j] = 1.0
        A[:, j] = apply_fn(e)
    return A


def eigen_cg(A, b, iters=10, tol=np.finfo(np.float32).eps):
    """Textbook Jacobi-preconditioned CG as in Eigen's conjugate_gradient()."""
    A = np.asarray(A, np.float64)
    b = np.asarray(b, np.float64)
    d = np.diag(A)
    inv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 1.0)
    x = np.zeros_like(b)
    r = b.copy()
    bn2 = b @ b
    if bn2 == 0:
        return x
    thr = max(tol * tol * bn2, np.finfo(np.float32).tiny)
    if r @ r < thr:
        return x
    p = inv * r
    absnew = r @ p
    for _ in range(iters):
        q = A @ p
        alpha = absnew / (p @ q)
        x += alpha * p
        r -= alpha * q
        if r @ r < thr:
            break
        z = inv * r
        absold = absnew
        absnew = r @ z
        p = z + (absnew / absold) * p
    return x
