"""ORACLE - TEST INFRASTRUCTURE ONLY.

ctypes wrapper of the CPU restatements of the reference's UAAMG-PCG Poisson
path (oracle/poisson_oracle.cpp) and variational viscosity path
(oracle/viscosity_oracle.cpp). Only tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py may import this package; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liborc.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    L = C.CDLL(str(LIB))
    vp, i32, i64, f32, f64 = C.c_void_p, C.c_int, C.c_int64, C.c_float, C.c_double
    sig = {
        "orc_set_threads": (None, [i32]),
        "orc_get_threads": (i32, []),
        "orc_fraction_inside2": (f32, [f32, f32]),
        "orc_fraction_inside4": (f32, [f32, f32, f32, f32]),
        "orc_poisson_create": (vp, [i64, vp, vp, vp, f32, i64, vp, vp, vp, vp, f32, f32, i64]),
        "orc_poisson_destroy": (None, [vp]),
        "orc_nlevels": (i32, [vp]),
        "orc_level_ndof": (i64, [vp, i32]),
        "orc_level_nleaf_tree": (i64, [vp, i32]),
        "orc_level_nleaf_dof": (i64, [vp, i32]),
        "orc_level_topology": (None, [vp, i32, vp, vp, vp]),
        "orc_level_coefficients": (None, [vp, i32, i32, vp, vp]),
        "orc_build_rhs": (None, [vp, i64, vp, vp, vp, i64, vp, vp, vp, vp, f32, i64, vp, vp, vp, vp]),
        "orc_apply": (i32, [vp, i32, i32, f32, vp, vp, vp]),
        "orc_face_weights": (None, [i64, vp, vp, i64, vp, vp, f32, f32, C.POINTER(i64), vp, vp, vp]),
        "orc_apply_pressure_gradient": (None, [vp, vp, i64, vp, vp, f32, i64, vp, vp, vp, i64, vp, vp, vp, i32,
                                               i64, vp, vp, vp, vp, vp, f32, f32]),
        "orc_first_sweep": (i32, [vp, i32, i32, f32, vp, vp]),
        "orc_restrict": (i32, [vp, i32, vp, vp]),
        "orc_prolong": (i32, [vp, i32, vp, vp, f32]),
        "orc_coarsest": (i32, [vp, vp, vp]),
        "orc_precondition": (i32, [vp, i32, vp, vp]),
        "orc_solve_pcg": (i32, [vp, i32, f32, i32, vp, vp, C.POINTER(i32), vp, i32]),
        "orc_solve_mg": (i32, [vp, i32, f32, i32, vp, vp, C.POINTER(i32)]),
        "orc_solve_pcg64": (i32, [vp, i32, f64, i32, vp, vp, C.POINTER(i32)]),
        "orc_apply64": (i32, [vp, i32, vp, vp]),
        "orc_volume_fraction8": (f32, [vp]),
        "orc_volume_fraction4": (f32, [f32, f32, f32, f32]),
        "orc_visc_create": (vp, [i64, vp, vp, vp, f32, i64, vp, vp, vp, f32, i64, vp, vp, vp, f32,
                                 f32, f32, f32, i64, i32]),
        "orc_visc_destroy": (None, [vp]),
        "orc_visc_nlevels": (i32, [vp]),
        "orc_visc_level_ndof": (i64, [vp, i32]),
        "orc_visc_level_nleaf": (i64, [vp, i32]),
        "orc_visc_level_default": (f32, [vp, i32, i32]),
        "orc_visc_level_topology": (None, [vp, i32, vp, vp, vp]),
        "orc_visc_level_coefficients": (None, [vp, i32, i32, vp, vp]),
        "orc_visc_level_active": (i64, [vp, i32, i32, vp, vp]),
        "orc_visc_face_volume": (None, [vp, i32, vp]),
        "orc_visc_volume_at": (None, [vp, i32, i64, vp, vp]),
        "orc_visc_dof_at": (None, [vp, i32, i32, i64, vp, vp]),
        "orc_visc_level_matrix": (i64, [vp, i32, vp, vp, vp]),
        "orc_visc_build_rhs": (None, [vp, i64, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp]),
        "orc_visc_apply": (i32, [vp, i32, i32, vp, vp, vp]),
        "orc_visc_restrict": (i32, [vp, i32, vp, vp]),
        "orc_visc_prolong": (i32, [vp, i32, vp, vp]),
        "orc_visc_coarsest": (i32, [vp, vp, vp]),
        "orc_visc_precondition": (i32, [vp, vp, vp]),
        "orc_visc_solve": (i32, [vp, f32, i32, vp, vp, vp]),
        "orc_visc_solve64": (i32, [vp, f64, i32, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OraclePoisson:
    """Reference-semantics CPU hierarchy + solver for one problem."""

    def __init__(self, phi, weights, dt: float, dx: float, max_coarsest: int = 4000):
        L = lib()
        self._keep = []

        def arr(a, dt_):
            a = np.ascontiguousarray(a, dtype=dt_)
            self._keep.append(a)
            return a

        po = arr(phi.origins, np.int32)
        pa = None if phi.active is None else arr(phi.active, np.uint64)
        pv = arr(phi.values, np.float32)
        wo = arr(weights.origins, np.int32)
        wa = None if weights.active is None else arr(weights.active, np.uint64)
        wv = arr(weights.values, np.float32)
        wbg = arr(np.array(weights.background, dtype=np.float32).reshape(-1)[:3], np.float32)
        phi_bg = float(np.array(phi.background).reshape(-1)[0])
        self.h = L.orc_poisson_create(po.shape[0], _p(po), _p(pa), _p(pv), phi_bg, wo.shape[0], _p(wo), _p(wa),
                                      _p(wv), _p(wbg), float(dt), float(dx), int(max_coarsest))
        self._keep = []
        self.dx = dx

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_poisson_destroy(self.h)
            self.h = None

    @property
    def nlevels(self) -> int:
        return lib().orc_nlevels(self.h)

    def ndof(self, level: int = 0) -> int:
        return lib().orc_level_ndof(self.h, level)

    def nleaf_tree(self, level: int) -> int:
        return lib().orc_level_nleaf_tree(self.h, level)

    def topology(self, level: int):
        n = lib().orc_level_nleaf_dof(self.h, level)
        org = np.zeros((n, 3), np.int32)
        mask = np.zeros((n, 8), np.uint64)
        base = np.zeros(n, np.int64)
        lib().orc_level_topology(self.h, level, _p(org), _p(mask), _p(base))
        return org, mask, base

    def coefficients(self, level: int, kind: int):
        n = lib().orc_level_nleaf_dof(self.h, level)
        br = np.zeros((n, 512), np.float32)
        kept = np.zeros(n, np.uint8)
        lib().orc_level_coefficients(self.h, level, kind, _p(br), _p(kept))
        return br, kept

    def build_rhs(self, weights, vx, vy, vz, solid_velocity=None) -> np.ndarray:
        """createPressurePoissonRightHandSide (simd_vdb_poisson.cpp:95-127). vx, vy, vz share one leaf
        topology; solid_velocity is a Vec3 grid (None: zero everywhere)."""
        rhs = np.zeros(self.ndof(0), np.float32)
        wo = np.ascontiguousarray(weights.origins, np.int32)
        wv = np.ascontiguousarray(weights.values, np.float32)
        wbg = np.array(weights.background, dtype=np.float32).reshape(-1)[:3].copy()
        vo = np.ascontiguousarray(vx.origins, np.int32)
        xs = [np.ascontiguousarray(v.values, np.float32) for v in (vx, vy, vz)]
        vbg = float(np.array(vx.background).reshape(-1)[0])
        if solid_velocity is None:
            so, sv, sbg = np.zeros((0, 3), np.int32), np.zeros((0, 512, 3), np.float32), np.zeros(3, np.float32)
        else:
            so = np.ascontiguousarray(solid_velocity.origins, np.int32)
            sv = np.ascontiguousarray(solid_velocity.values, np.float32)
            sbg = np.array(solid_velocity.background, dtype=np.float32).reshape(-1)
            sbg = np.resize(sbg, 3).astype(np.float32)
        lib().orc_build_rhs(self.h, wo.shape[0], _p(wo), _p(wv), _p(wbg), vo.shape[0], _p(vo), *[_p(x) for x in xs],
                            vbg, so.shape[0], _p(so), _p(sv), _p(sbg), _p(rhs))
        return rhs

    def apply_pressure_gradient(self, pressure, phi, weights, solid, velocity, dt: float, friction: float):
        """FLIP_Solver_OpenVDB::apply_pressure_gradient restated (FLIP_vdb.cpp:2571-2667).
        velocity: 3 (origins [n,3], active [n,8], values [n,512]) per channel; returns per channel
        (active, values, update, reflected) after the update."""
        pr = np.ascontiguousarray(pressure, np.float32)
        pa = np.ascontiguousarray(phi.origins, np.int32)
        pv = np.ascontiguousarray(phi.values, np.float32)
        wa = np.ascontiguousarray(weights.origins, np.int32)
        wv = np.ascontiguousarray(weights.values, np.float32)
        wb = np.asarray(weights.background, np.float32).reshape(-1)
        if solid is None:
            sa, sv, sb = np.zeros((0, 3), np.int32), np.zeros((0, 512, 3), np.float32), np.zeros(3, np.float32)
        else:
            sa = np.ascontiguousarray(solid.origins, np.int32)
            sv = np.ascontiguousarray(solid.values, np.float32)
            sb = np.asarray(solid.background, np.float32).reshape(-1)
        out = []
        for c, (org, act, val) in enumerate(velocity):
            org = np.ascontiguousarray(org, np.int32)
            act = np.array(act, dtype=np.uint64, copy=True)
            val = np.array(val, dtype=np.float32, copy=True)
            upd = np.zeros_like(val)
            refl = np.zeros_like(val)
            bg = phi.background[0] if isinstance(phi.background, (tuple, list)) else phi.background
            lib().orc_apply_pressure_gradient(self.h, _p(pr), pa.shape[0], _p(pa), _p(pv), float(bg), wa.shape[0],
                                              _p(wa), _p(wv), _p(wb), sa.shape[0], _p(sa), _p(sv), _p(sb), c,
                                              org.shape[0], _p(org), _p(act), _p(val), _p(upd), _p(refl),
                                              float(dt), float(friction))
            out.append((act, val, upd, refl))
        return out

    def apply(self, level, mode, x, b=None, weight=0.0):
        x = np.ascontiguousarray(x, np.float32)
        b = None if b is None else np.ascontiguousarray(b, np.float32)
        y = np.zeros(self.ndof(level), np.float32)
        lib().orc_apply(self.h, level, mode, float(weight), _p(x), _p(b), _p(y))
        return y

    def first_sweep(self, level, kind, b, weight):
        b = np.ascontiguousarray(b, np.float32)
        y = np.zeros(self.ndof(level), np.float32)
        lib().orc_first_sweep(self.h, level, kind, float(weight), _p(b), _p(y))
        return y

    def restriction(self, fine_level, fine):
        fine = np.ascontiguousarray(fine, np.float32)
        out = np.zeros(self.ndof(fine_level + 1), np.float32)
        lib().orc_restrict(self.h, fine_level, _p(fine), _p(out))
        return out

    def prolongation(self, fine_level, coarse, fine, alpha=1.0):
        coarse = np.ascontiguousarray(coarse, np.float32)
        out = np.array(fine, np.float32, copy=True)
        lib().orc_prolong(self.h, fine_level, _p(coarse), _p(out), float(alpha))
        return out

    def coarsest_solve(self, b):
        b = np.ascontiguousarray(b, np.float32)
        x = np.zeros(self.ndof(self.nlevels - 1), np.float32)
        lib().orc_coarsest(self.h, _p(b), _p(x))
        return x

    def precondition(self, r, smoother=0):
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros_like(r)
        lib().orc_precondition(self.h, smoother, _p(r), _p(z))
        return z

    def solve_pcg(self, b, x0=None, tol=1e-7, maxit=100, smoother=0, history=False):
        x = np.zeros(self.ndof(0), np.float32) if x0 is None else np.array(x0, np.float32, copy=True)
        b = np.ascontiguousarray(b, np.float32)
        it = C.c_int()
        hist = np.zeros(maxit + 2, np.float32) if history else None
        st = lib().orc_solve_pcg(self.h, smoother, float(tol), maxit, _p(x), _p(b), C.byref(it), _p(hist),
                                 0 if hist is None else hist.size)
        if history:
            return st, it.value, x, hist[: it.value + 2]
        return st, it.value, x

    def solve_mg(self, b, x0=None, tol=1e-7, maxit=100, smoother=0):
        x = np.zeros(self.ndof(0), np.float32) if x0 is None else np.array(x0, np.float32, copy=True)
        b = np.ascontiguousarray(b, np.float32)
        it = C.c_int()
        st = lib().orc_solve_mg(self.h, smoother, float(tol), maxit, _p(x), _p(b), C.byref(it))
        return st, it.value, x

    def solve_pcg64(self, b, x0=None, tol=1e-12, maxit=200, smoother=0):
        x = np.zeros(self.ndof(0), np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
        b = np.ascontiguousarray(b, np.float64)
        it = C.c_int()
        st = lib().orc_solve_pcg64(self.h, smoother, float(tol), maxit, _p(x), _p(b), C.byref(it))
        return st, it.value, x

    def apply64(self, x, level=0):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.ndof(level), np.float64)
        lib().orc_apply64(self.h, level, _p(x), _p(y))
        return y


def face_weights(liquid, solid, cut: bool = False):
    """FLIP_Solver_OpenVDB::calculate_face_weights restated (FLIP_vdb.cpp:1990-2068; cut=True: the
    member variant of FLIP_vdb.cpp:2669-2752, weights below 1e-2 set to 0): (origins, active,
    values) of the weight grid, leaves in tree order."""
    lo = np.ascontiguousarray(liquid.origins, np.int32)
    la = (np.full((lo.shape[0], 8), np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64) if liquid.active is None
          else np.ascontiguousarray(liquid.active, np.uint64))
    so = np.ascontiguousarray(solid.origins, np.int32)
    sv = np.ascontiguousarray(solid.values, np.float32)
    bg = solid.background[0] if isinstance(solid.background, (tuple, list)) else solid.background
    n = C.c_int64()
    c = C.c_float(1e-2 if cut else 0.0)
    lib().orc_face_weights(lo.shape[0], _p(lo), _p(la), so.shape[0], _p(so), _p(sv), float(bg), c, C.byref(n),
                           None, None, None)
    org = np.zeros((n.value, 3), np.int32)
    act = np.zeros((n.value, 8), np.uint64)
    val = np.zeros((n.value, 512, 3), np.float32)
    lib().orc_face_weights(lo.shape[0], _p(lo), _p(la), so.shape[0], _p(so), _p(sv), float(bg), c, C.byref(n),
                           _p(org), _p(act), _p(val))
    return org, act, val


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return lib().orc_get_threads()


def fraction_inside2(a, b):
    return lib().orc_fraction_inside2(a, b)


def fraction_inside4(bl, br, tl, tr):
    return lib().orc_fraction_inside4(bl, br, tl, tr)


def volume_fraction8(p) -> float:
    a = np.ascontiguousarray(p, dtype=np.float32)
    return float(lib().orc_volume_fraction8(_p(a)))


def volume_fraction4(a, b, c, d) -> float:
    return float(lib().orc_volume_fraction4(a, b, c, d))


class OracleViscosity:
    """Reference-semantics CPU viscosity hierarchy + solver (simd_viscosity3d)."""

    def __init__(self, viscosity, liquid, solid, dt: float, rho: float, dx: float, max_coarsest: int = 4000,
                 trim: bool = True):
        L = lib()
        self._keep = []

        def g(grid):
            o = np.ascontiguousarray(grid.origins, np.int32)
            v = np.ascontiguousarray(grid.values, np.float32)
            a = None if grid.active is None else np.ascontiguousarray(grid.active, np.uint64)
            self._keep += [o, v, a]
            bg = grid.background[0] if isinstance(grid.background, (tuple, list)) else grid.background
            return [o.shape[0], _p(o), _p(a), _p(v), float(bg)]

        self._h = L.orc_visc_create(*g(liquid), *g(solid), *g(viscosity), float(dt), float(rho), float(dx),
                                    int(max_coarsest), int(trim))
        self._keep = []
        self.nlevels = int(L.orc_visc_nlevels(self._h))
        self.ndof = self.level_ndof(0)

    def level_ndof(self, level: int) -> int:
        return int(lib().orc_visc_level_ndof(self._h, level))

    def level_nleaf(self, level: int) -> int:
        return int(lib().orc_visc_level_nleaf(self._h, level))

    def defaults(self, level: int):
        return float(lib().orc_visc_level_default(self._h, level, 0)), float(lib().orc_visc_level_default(
            self._h, level, 1))

    def topology(self, level: int):
        n = self.level_nleaf(level)
        org = np.zeros((n, 3), np.int32)
        mask = np.zeros((n, 3, 8), np.uint64)
        base = np.zeros((n, 3), np.int64)
        lib().orc_visc_level_topology(self._h, level, _p(org), _p(mask), _p(base))
        return org, mask, base

    def coefficients(self, level: int, grid: int):
        n = self.level_nleaf(level)
        br = np.zeros((n, 512), np.float32)
        kept = np.zeros(n, np.uint8)
        lib().orc_visc_level_coefficients(self._h, level, grid, _p(br), _p(kept))
        return br, kept

    def active(self, level: int, grid: int):
        n = int(lib().orc_visc_level_active(self._h, level, grid, None, None))
        xyz = np.zeros((n, 3), np.int32)
        val = np.zeros(n, np.float32)
        lib().orc_visc_level_active(self._h, level, grid, _p(xyz), _p(val))
        return xyz, val

    def face_volume(self, c: int):
        out = np.zeros(self.ndof, np.float32)
        lib().orc_visc_face_volume(self._h, c, _p(out))
        return out

    def volume_at(self, kind: int, xyz):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        out = np.zeros(xyz.shape[0], np.float32)
        lib().orc_visc_volume_at(self._h, kind, xyz.shape[0], _p(xyz), _p(out))
        return out

    def dof_at(self, level: int, c: int, xyz):
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        out = np.zeros(xyz.shape[0], np.int32)
        lib().orc_visc_dof_at(self._h, level, c, xyz.shape[0], _p(xyz), _p(out))
        return out

    def matrix(self, level: int):
        nnz = int(lib().orc_visc_level_matrix(self._h, level, None, None, None))
        n = self.level_ndof(level)
        ptr = np.zeros(n + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz, np.float32)
        lib().orc_visc_level_matrix(self._h, level, _p(ptr), _p(col), _p(val))
        return ptr, col, val

    def build_rhs(self, velocity, solid_velocity=None):
        keep = []

        def g(grid, default_n=0):
            if grid is None:
                return [0, None, None, None, None]
            o = np.ascontiguousarray(grid.origins, np.int32)
            v = np.ascontiguousarray(grid.values, np.float32)
            a = None if grid.active is None else np.ascontiguousarray(grid.active, np.uint64)
            bg = np.asarray(grid.background, np.float32).reshape(-1)
            bg = np.ascontiguousarray(np.resize(bg, 3), np.float32)
            keep.extend([o, v, a, bg])
            return [o.shape[0], _p(o), _p(a), _p(v), _p(bg)]

        out = np.zeros(self.ndof, np.float32)
        lib().orc_visc_build_rhs(self._h, *g(velocity), *g(solid_velocity), _p(out))
        return out

    def apply(self, level: int, mode: int, x, b=None):
        x = np.ascontiguousarray(x, np.float32)
        b = np.zeros_like(x) if b is None else np.ascontiguousarray(b, np.float32)
        y = np.zeros(self.level_ndof(level), np.float32)
        lib().orc_visc_apply(self._h, level, mode, _p(x), _p(b), _p(y))
        return y

    def restriction(self, fine_level: int, fine):
        fine = np.ascontiguousarray(fine, np.float32)
        out = np.zeros(self.level_ndof(fine_level + 1), np.float32)
        lib().orc_visc_restrict(self._h, fine_level, _p(fine), _p(out))
        return out

    def prolongation(self, fine_level: int, coarse, fine):
        coarse = np.ascontiguousarray(coarse, np.float32)
        out = np.array(fine, np.float32, copy=True)
        lib().orc_visc_prolong(self._h, fine_level, _p(coarse), _p(out))
        return out

    def coarsest_solve(self, b):
        b = np.ascontiguousarray(b, np.float32)
        x = np.zeros(self.level_ndof(self.nlevels - 1), np.float32)
        lib().orc_visc_coarsest(self._h, _p(b), _p(x))
        return x

    def precondition(self, r):
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros_like(r)
        lib().orc_visc_precondition(self._h, _p(r), _p(z))
        return z

    def solve(self, b, x0=None, tol: float = 1e-6, maxit: int = 100):
        """pcg_solve; returns (m_iteration, x, ||r||_inf history)."""
        x = np.zeros(self.ndof, np.float32) if x0 is None else np.array(x0, np.float32, copy=True)
        b = np.ascontiguousarray(b, np.float32)
        hist = np.zeros(maxit + 2, np.float32)
        it = int(lib().orc_visc_solve(self._h, float(tol), int(maxit), _p(x), _p(b), _p(hist)))
        return it, x, hist[: it + 2]

    def solve64(self, b, x0=None, tol: float = 1e-6, maxit: int = 100):
        x = np.zeros(self.ndof, np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
        b = np.ascontiguousarray(b, np.float64)
        it = int(lib().orc_visc_solve64(self._h, float(tol), int(maxit), _p(x), _p(b)))
        return it, x

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_visc_destroy(h)
            self._h = None
