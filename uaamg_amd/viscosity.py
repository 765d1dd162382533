"""Variational viscosity solve: Python mirror of simd_uaamg::L_with_level and
simd_uaamg::simd_viscosity3d (SIMD_UAAMG/simd_viscosity3d.h:12-321).

Build the level-0 viscosity operator from the viscosity, liquid-SDF and
solid-SDF grids (L_with_level.create_viscosity_matrix), build the right-hand
side from the face velocities (build_rhs), then simd_viscosity3d.pcg_solve
(in_lhs, tolerance). Velocities are DOF-ordered float32 numpy arrays
(channel 0's DOFs, then channel 1's, then channel 2's; packed_FloatGrid3 is
that vector plus the channel split). The work runs through the C ABI of
libuaamg_amd.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Grid, _check, _Grid, _ptr, _ViscOptions, lib

# light_weight_applier::working_mode (simd_viscosity3d.h:13) and the 6-colour sweeps
NORMAL, RESIDUAL, JACOBI, SPAI0, RED_GS, BLACK_GS, XYZ_RBGS, ZYX_RBGS = range(8)


def grid_diag(c: int) -> int:
    return c


def grid_invdiag(c: int) -> int:
    return 3 + c


def grid_neg(c: int, axis: int) -> int:
    return 6 + 3 * c + axis


def grid_cross(normal: int, j: int, k: int) -> int:
    return 15 + 4 * normal + 2 * j + k


def grid_spai0(c: int) -> int:
    return 27 + c


class packed_FloatGrid3:
    """Three face-velocity channels as one DOF-ordered vector (packed_FloatGrid3, packed3grids.h:5-23)."""

    def __init__(self, data: np.ndarray, nchan: tuple):
        self.data = data
        self.nchan = tuple(int(n) for n in nchan)

    @property
    def v(self) -> list:
        a, b, _ = self.nchan
        return [self.data[:a], self.data[a:a + b], self.data[a + b:]]

    def deepCopy(self) -> "packed_FloatGrid3":
        return packed_FloatGrid3(self.data.copy(), self.nchan)


class ViscScene:
    """Synthetic viscosity scene built by the library (liquid / solid SDFs, viscosity, velocities)."""

    def __init__(self, config: int = 0, n: int = 0, seed: int = 0, mu: float = 1.0, dt: float = 1.0 / 24.0,
                 rho: float = 1000.0):
        L = lib()
        h = C.c_void_p()
        _check(L.uaamg_visc_scene_create(int(config), int(n), C.c_uint64(seed), float(mu), float(dt), float(rho),
                                         C.byref(h)), "visc_scene_create")
        self._h = h
        gs = [_Grid() for _ in range(4)]
        dx, dtv, rhov = C.c_float(), C.c_float(), C.c_float()
        _check(L.uaamg_visc_scene_grids(h, *[C.byref(g) for g in gs], C.byref(dx), C.byref(dtv), C.byref(rhov)),
               "visc_scene_grids")
        self.dx, self.dt, self.rho, self.mu = float(dx.value), float(dtv.value), float(rhov.value), float(mu)
        self.liquid, self.solid, self.viscosity, self.velocity = (self._wrap(g) for g in gs)

    @staticmethod
    def _wrap(g: _Grid) -> Grid:
        n = g.nleaf
        if n == 0:
            return Grid(np.zeros((0, 3), np.int32), np.zeros((0, 512) if g.ncomp == 1 else (0, 512, 3), np.float32),
                        None, tuple(g.background))
        org = np.ctypeslib.as_array(C.cast(g.origins, C.POINTER(C.c_int32)), shape=(n * 3,))
        shape = (n, 512) if g.ncomp == 1 else (n, 512, 3)
        val = np.ctypeslib.as_array(C.cast(g.values, C.POINTER(C.c_float)), shape=(n * 512 * g.ncomp,))
        act = None
        if g.active:
            act = np.ctypeslib.as_array(C.cast(g.active, C.POINTER(C.c_uint64)), shape=(n * 8,)).reshape(n, 8).copy()
        return Grid(org.reshape(n, 3).copy(), val.reshape(shape).copy(), act, tuple(g.background))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().uaamg_visc_scene_destroy(h)
            self._h = None


class L_with_level:
    """Level-0 viscosity operator plus its coarse levels (L_with_level, simd_viscosity3d.h:12-281).

    create_viscosity_matrix(viscosity, liquid_sdf, solid_sdf, dt, rho, dx) mirrors
    simd_viscosity3d.h:95-98; dx is liquid_sdf's voxel size. device=-1 keeps a
    host-only handle (inspection; no solves). precision=1 stores the W-cycle's
    level vectors below level 0 in BF16 (FP32 arithmetic; no reference
    counterpart).
    """

    def __init__(self, viscosity: Grid, liquid_sdf: Grid, solid_sdf: Grid, dt: float, rho: float, dx: float, *,
                 device: int = 0, max_coarsest_dof: int = 4000, flags: int = 0, precision: int = 0):
        L = lib()
        o = _ViscOptions()
        L.uaamg_viscosity_options_default(C.byref(o))
        o.device = device
        o.max_coarsest_dof = max_coarsest_dof
        o.flags = flags
        o.precision = precision
        h = C.c_void_p()
        g = [viscosity._struct(), liquid_sdf._struct(), solid_sdf._struct()]
        _check(L.uaamg_viscosity_create(*[C.byref(x) for x in g], float(dt), float(rho), float(dx), C.byref(o),
                                        C.byref(h)), "uaamg_viscosity_create")
        self._h = h
        self.m_dt, self.m_rho, self.m_dx = float(dt), float(rho), float(dx)
        self.m_level = 0
        self.m_ndof = int(L.uaamg_viscosity_ndof(h))

    @classmethod
    def create_viscosity_matrix(cls, viscosity: Grid, liquid_sdf: Grid, solid_sdf: Grid, dt: float, rho: float,
                                dx: float, **kw) -> "L_with_level":
        return cls(viscosity, liquid_sdf, solid_sdf, dt, rho, dx, **kw)

    @property
    def nlevels(self) -> int:
        return int(lib().uaamg_viscosity_nlevels(self._h))

    def kernel_map(self) -> list[dict]:
        """Per level: whether half-visits run as single-workgroup launches, and the timings behind it."""
        out = []
        for l in range(self.nlevels):
            b, ms, mp = C.c_int32(), C.c_float(), C.c_float()
            _check(lib().uaamg_viscosity_kernel_map(self._h, l, C.byref(b), C.byref(ms), C.byref(mp)), "kernel_map")
            out.append({"level": l, "small": bool(b.value), "ms_small": round(ms.value, 4),
                        "ms_pass": round(mp.value, 4)})
        return out

    def level_info(self, level: int) -> dict:
        nd, nl = C.c_int64(), C.c_int64()
        ch = np.zeros(3, np.int64)
        _check(lib().uaamg_viscosity_level_info(self._h, level, C.byref(nd), _ptr(ch), C.byref(nl)), "level_info")
        return {"ndof": nd.value, "nchan": tuple(int(v) for v in ch), "nleaf_dof": nl.value}

    def level_ndof(self, level: int) -> int:
        return self.level_info(level)["ndof"]

    def level_topology(self, level: int):
        n = self.level_info(level)["nleaf_dof"]
        org = np.zeros((n, 3), np.int32)
        mask = np.zeros((n, 3, 8), np.uint64)
        base = np.zeros((n, 3), np.int64)
        _check(lib().uaamg_viscosity_level_topology(self._h, level, _ptr(org), _ptr(mask), _ptr(base)), "topology")
        return org, mask, base

    def level_coefficients(self, level: int, grid: int):
        n = self.level_info(level)["nleaf_dof"]
        br = np.zeros((n, 512), np.float32)
        kept = np.zeros(n, np.uint8)
        _check(lib().uaamg_viscosity_level_coefficients(self._h, level, grid, _ptr(br), _ptr(kept)), "coefficients")
        return br, kept

    def level_matrix(self, level: int):
        """(ptr, col, val) CSR of the level's explicit matrix (get_triplets)."""
        nnz = C.c_int64()
        _check(lib().uaamg_viscosity_level_matrix(self._h, level, C.byref(nnz), None, None, None), "matrix")
        n = self.level_ndof(level)
        ptr = np.zeros(n + 1, np.int64)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value, np.float32)
        _check(lib().uaamg_viscosity_level_matrix(self._h, level, C.byref(nnz), _ptr(ptr), _ptr(col), _ptr(val)),
               "matrix")
        return ptr, col, val

    def get_zero_vec(self) -> packed_FloatGrid3:
        return packed_FloatGrid3(np.zeros(self.m_ndof, np.float32), self.level_info(0)["nchan"])

    def build_rhs(self, velocity: Grid, solid_velocity: Grid | None = None) -> packed_FloatGrid3:
        out = np.zeros(self.m_ndof, np.float32)
        gu = velocity._struct()
        gs = None if solid_velocity is None else C.byref(solid_velocity._struct())
        _check(lib().uaamg_viscosity_build_rhs(self._h, C.byref(gu), gs, _ptr(out)), "build_rhs")
        return packed_FloatGrid3(out, self.level_info(0)["nchan"])

    # ---- level operations on DOF-ordered vectors
    def apply(self, level: int, mode: int, x: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        b = None if b is None else np.ascontiguousarray(b, np.float32)
        y = np.zeros(self.level_ndof(level), np.float32)
        _check(lib().uaamg_viscosity_apply(self._h, level, mode, _ptr(x), _ptr(b), _ptr(y)), "apply")
        return y

    def L_apply(self, x, level: int = 0):
        return self.apply(level, NORMAL, x)

    def residual_apply(self, x, b, level: int = 0):
        return self.apply(level, RESIDUAL, x, b)

    def Jacobi_apply(self, x, b, level: int = 0):
        return self.apply(level, JACOBI, x, b)

    def SPAI0_apply(self, x, b, level: int = 0):
        return self.apply(level, SPAI0, x, b)

    def XYZ_RBGS_apply(self, x, b, level: int = 0):
        return self.apply(level, XYZ_RBGS, x, b)

    def ZYX_RBGS_apply(self, x, b, level: int = 0):
        return self.apply(level, ZYX_RBGS, x, b)

    def restriction(self, fine_level: int, fine: np.ndarray) -> np.ndarray:
        fine = np.ascontiguousarray(fine, np.float32)
        out = np.zeros(self.level_ndof(fine_level + 1), np.float32)
        _check(lib().uaamg_viscosity_restrict(self._h, fine_level, _ptr(fine), _ptr(out)), "restrict")
        return out

    def prolongation(self, fine_level: int, coarse: np.ndarray, fine: np.ndarray) -> np.ndarray:
        coarse = np.ascontiguousarray(coarse, np.float32)
        out = np.array(fine, dtype=np.float32, copy=True)
        _check(lib().uaamg_viscosity_prolong(self._h, fine_level, _ptr(coarse), _ptr(out)), "prolong")
        return out

    def solve_exact(self, b: np.ndarray) -> np.ndarray:
        b = np.ascontiguousarray(b, np.float32)
        x = np.zeros(self.level_ndof(self.nlevels - 1), np.float32)
        _check(lib().uaamg_viscosity_coarsest_solve(self._h, _ptr(b), _ptr(x)), "solve_exact")
        return x

    def algorithmic_bytes(self, level: int, mode: int, channel: int) -> float:
        v = C.c_double()
        _check(lib().uaamg_viscosity_algorithmic_bytes(self._h, level, mode, channel, C.byref(v)), "bytes")
        return float(v.value)

    def time_apply(self, level: int, mode: int, channel: int, reps: int = 50) -> float:
        v = C.c_float()
        _check(lib().uaamg_viscosity_time_apply(self._h, level, mode, channel, reps, C.byref(v)), "time_apply")
        return float(v.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().uaamg_viscosity_destroy(h)
            self._h = None


class simd_viscosity3d:
    """simd_viscosity3d(level0, velocity, solid_velocity) (simd_viscosity3d.h:292-294)."""

    def __init__(self, level0: L_with_level, in_liquid_velocity: Grid, in_solid_velocity: Grid | None = None):
        self.m_matrix_levels = level0
        self.m_rhs = level0.build_rhs(in_liquid_velocity, in_solid_velocity)
        self.m_iteration = 0
        self.m_max_iter = 100

    def _push(self):
        o = _ViscOptions()
        lib().uaamg_viscosity_options_default(C.byref(o))
        o.max_iterations = int(self.m_max_iter)
        _check(lib().uaamg_viscosity_set_options(self.m_matrix_levels._h, C.byref(o)), "set_options")

    def pcg_solve(self, in_lhs, tolerance: float, rhs: np.ndarray | None = None) -> None:
        """In place on in_lhs (packed_FloatGrid3 or float32 array of length m_ndof)."""
        self._push()
        x = in_lhs.data if isinstance(in_lhs, packed_FloatGrid3) else in_lhs
        if x.dtype != np.float32 or not x.flags.c_contiguous:
            raise TypeError("in_lhs must be a contiguous float32 array")
        b = self.m_rhs.data if rhs is None else np.ascontiguousarray(rhs, np.float32)
        it = C.c_int32()
        _check(lib().uaamg_viscosity_solve_pcg(self.m_matrix_levels._h, float(tolerance), _ptr(x), _ptr(b),
                                               C.byref(it)), "pcg_solve")
        self.m_iteration = int(it.value)

    def solve_pcg_device(self, tolerance: float, x_ptr: int, b_ptr: int) -> int:
        """pcg_solve on DOF-ordered device vectors given as raw device pointers; returns m_iteration."""
        self._push()
        it = C.c_int32()
        _check(lib().uaamg_viscosity_solve_pcg_device(self.m_matrix_levels._h, float(tolerance), C.c_void_p(x_ptr),
                                                      C.c_void_p(b_ptr), C.byref(it)), "pcg_solve_device")
        self.m_iteration = int(it.value)
        return self.m_iteration

    def precondition(self, r: np.ndarray) -> np.ndarray:
        """One W-cycle from zero, z = M r (mucycle(z, r, 2, true, 0, 2))."""
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros_like(r)
        _check(lib().uaamg_viscosity_precondition(self.m_matrix_levels._h, _ptr(r), _ptr(z)), "precondition")
        return z


class ViscBatch:
    """Many viscosity problems merged into one leaf table per level and solved
    together (uaamg_viscosity_batch_*): one launch per colour pass, transfer and
    PCG vector operation for the whole batch, per-problem PCG scalars and
    stopping rule on the device; each problem's result and m_iteration equal
    its pcg_solve alone (simd_viscosity3d.cpp:3694-3740). `problems` is a list
    of (viscosity, liquid_sdf, solid_sdf, dt, rho, dx)."""

    def __init__(self, problems, *, max_iterations: int = 100, device: int = 0, max_coarsest_dof: int = 4000,
                 flags: int = 0):
        L = lib()
        n = len(problems)
        o = _ViscOptions()
        L.uaamg_viscosity_options_default(C.byref(o))
        o.device, o.max_coarsest_dof, o.flags = device, max_coarsest_dof, flags
        o.max_iterations = int(max_iterations)
        self._grids = [(p[0]._struct(), p[1]._struct(), p[2]._struct()) for p in problems]
        arrs = [(C.c_void_p * n)(*[C.addressof(g[k]) for g in self._grids]) for k in range(3)]
        dts = np.array([p[3] for p in problems], np.float32)
        rhos = np.array([p[4] for p in problems], np.float32)
        dxs = np.array([p[5] for p in problems], np.float32)
        h = C.c_void_p()
        _check(L.uaamg_viscosity_batch_create(n, arrs[0], arrs[1], arrs[2], _ptr(dts), _ptr(rhos), _ptr(dxs),
                                              C.byref(o), C.byref(h)), "uaamg_viscosity_batch_create")
        self._h = h
        self.n = n

    def info(self) -> dict:
        n, nl = C.c_int32(), C.c_int32()
        l0, nb = C.c_int64(), C.c_int64()
        _check(lib().uaamg_viscosity_batch_info(self._h, C.byref(n), C.byref(nl), C.byref(l0), C.byref(nb)),
               "info")
        return {"nproblems": n.value, "nlevels": nl.value, "leaves0": l0.value, "device_bytes": nb.value}

    def ndof(self, i: int) -> int:
        return int(lib().uaamg_viscosity_batch_ndof(self._h, i))

    def solve_device(self, tolerance: float, x_ptrs, b_ptrs) -> list[int]:
        """x_ptrs / b_ptrs: device addresses of each problem's DOF vectors (x: initial guess in, result
        out); returns each problem's m_iteration."""
        xs = (C.c_void_p * self.n)(*x_ptrs)
        bs = (C.c_void_p * self.n)(*b_ptrs)
        it = np.zeros(self.n, np.int32)
        _check(lib().uaamg_viscosity_batch_solve_device(self._h, float(tolerance), xs, bs, _ptr(it)),
               "viscosity batch solve")
        return it.tolist()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().uaamg_viscosity_batch_destroy(h)
            self._h = None
