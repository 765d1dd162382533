"""MI355X UAAMG-PCG pressure-Poisson solver: Python mirror of the reference API.

The classes mirror simd_uaamg::LaplacianWithLevel and simd_uaamg::PoissonSolver
(SIMD_UAAMG/simd_vdb_poisson.h:13-207): construct the Laplacian from a liquid
SDF grid and a face-weight grid, build the right-hand side, then
solveMultigridPCG(pressure, rhs). Grids are flattened OpenVDB grids (see
``Grid``); vectors are numpy arrays of length mNumDof in DOF order.

Everything runs through the C ABI of libuaamg_amd.so (include/uaamg_capi.h),
whose kernels are hand-written HIP for gfx950. There is no CPU fallback: if the
library is missing or no GPU is visible, the solver classes raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

__all__ = [
    "Grid", "Scene", "LaplacianWithLevel", "PoissonSolver", "SmootherOption", "library_path", "lib",
    "UaamgError", "SUCCESS", "FAILED", "device_count", "L_with_level", "simd_viscosity3d", "ViscScene",
    "packed_FloatGrid3", "solve_batch_device", "PoissonBatch", "ViscBatch",
]

_PKG = Path(__file__).resolve().parent
_LIBNAME = "libuaamg_amd.so"

SUCCESS = 0
FAILED = 1


class UaamgError(RuntimeError):
    pass


class _Grid(C.Structure):
    _fields_ = [
        ("nleaf", C.c_int64),
        ("origins", C.c_void_p),
        ("active", C.c_void_p),
        ("values", C.c_void_p),
        ("ncomp", C.c_int32),
        ("background", C.c_float * 3),
    ]


class _FaceChannel(C.Structure):
    _fields_ = [
        ("nleaf", C.c_int64),
        ("origins", C.c_void_p),
        ("active", C.c_void_p),
        ("values", C.c_void_p),
        ("update", C.c_void_p),
        ("reflected", C.c_void_p),
    ]


class _Options(C.Structure):
    _fields_ = [
        ("relative_tolerance", C.c_float),
        ("max_iterations", C.c_int32),
        ("smoother", C.c_int32),
        ("precision", C.c_int32),
        ("device", C.c_int32),
        ("max_coarsest_dof", C.c_int32),
        ("flags", C.c_int32),
        ("stream", C.c_void_p),
    ]


class _ViscOptions(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32),
        ("device", C.c_int32),
        ("max_coarsest_dof", C.c_int32),
        ("flags", C.c_int32),
        ("precision", C.c_int32),
        ("stream", C.c_void_p),
    ]


def library_path() -> Path:
    env = os.environ.get("UAAMG_LIBRARY")
    return Path(env) if env else _PKG / _LIBNAME


_lib_handle = None


def lib() -> C.CDLL:
    """Loads libuaamg_amd.so (built in-tree by uaamg_amd.build); raises if absent."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    path = library_path()
    if not path.exists():
        raise UaamgError(f"{path} not built: run `python -m uaamg_amd.build` (HIP extension missing)")
    # One HIP runtime per process: PyTorch ships its own, under the same soname.
    # Loaded first, it is the one this library binds to; loaded after this
    # library, it would come in beside the system runtime and see no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(path))
    i32, i64, f32, vp = C.c_int32, C.c_int64, C.c_float, C.c_void_p
    P = C.POINTER
    sig = {
        "uaamg_options_default": (C.c_int, [P(_Options)]),
        "uaamg_last_error": (C.c_char_p, []),
        "uaamg_version": (C.c_char_p, []),
        "uaamg_device_count": (C.c_int, []),
        "uaamg_poisson_create": (C.c_int, [P(_Grid), P(_Grid), f32, f32, P(_Options), P(vp)]),
        "uaamg_poisson_destroy": (None, [vp]),
        "uaamg_poisson_set_options": (C.c_int, [vp, P(_Options)]),
        "uaamg_poisson_ndof": (i64, [vp]),
        "uaamg_poisson_nlevels": (C.c_int, [vp]),
        "uaamg_poisson_level_info": (C.c_int, [vp, C.c_int, P(i64), P(i64), P(i64)]),
        "uaamg_poisson_level_topology": (C.c_int, [vp, C.c_int, vp, vp, vp]),
        "uaamg_poisson_grid_to_vector": (C.c_int, [vp, P(_Grid), vp]),
        "uaamg_poisson_vector_to_grid": (C.c_int, [vp, vp, vp]),
        "uaamg_poisson_level_coefficients": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
        "uaamg_poisson_build_rhs": (C.c_int, [vp, P(_Grid), P(_Grid), P(_Grid), P(_Grid), P(_Grid), vp]),
        "uaamg_poisson_build_rhs_device": (C.c_int, [vp, P(_Grid), P(_Grid), P(_Grid), P(_Grid), P(_Grid), i32, vp]),
        "uaamg_poisson_build_rhs_device_noalloc": (C.c_int, [vp, P(_Grid), P(_Grid), P(_Grid), P(_Grid), P(_Grid),
                                                           vp]),
        "uaamg_face_weights": (C.c_int, [P(_Grid), P(_Grid), i32, P(i64), vp, vp, vp]),
        "uaamg_face_weights_ex": (C.c_int, [P(_Grid), P(_Grid), i32, i32, P(i64), vp, vp, vp]),
        "uaamg_poisson_apply_pressure_gradient_device": (C.c_int, [vp, vp, P(_Grid), P(_Grid), P(_Grid),
                                                                   P(_FaceChannel), f32, f32, i32]),
        "uaamg_poisson_solve_pcg": (C.c_int, [vp, vp, vp, P(i32)]),
        "uaamg_poisson_solve_pcg_device": (C.c_int, [vp, vp, vp, P(i32)]),
        "uaamg_poisson_last_residual": (C.c_int, [vp, P(f32), P(i32)]),
        "uaamg_poisson_kernel_map": (C.c_int, [vp, C.c_int, P(i32), P(f32), P(f32)]),
        "uaamg_poisson_level_triplets": (C.c_int, [vp, C.c_int, P(i64), vp, vp, vp]),
        "uaamg_poisson_solve_mg": (C.c_int, [vp, vp, vp, P(i32)]),
        "uaamg_poisson_solve_batch_device": (C.c_int, [vp, i32, vp, vp, vp, vp, i32]),
        "uaamg_poisson_batch_create": (C.c_int, [i32, vp, vp, vp, vp, P(_Options), P(vp)]),
        "uaamg_poisson_batch_destroy": (None, [vp]),
        "uaamg_poisson_batch_set_options": (C.c_int, [vp, P(_Options)]),
        "uaamg_poisson_batch_info": (C.c_int, [vp, P(i32), P(i32), P(i64), P(i64)]),
        "uaamg_poisson_batch_ndof": (i64, [vp, i32]),
        "uaamg_poisson_batch_solve_device": (C.c_int, [vp, vp, vp, vp, vp]),
        "uaamg_poisson_batch_last_residual": (C.c_int, [vp, vp, vp]),
        "uaamg_poisson_batch_last_fallback": (C.c_int, [vp, vp, vp]),
        "uaamg_viscosity_solve_batch_device": (C.c_int, [vp, i32, f32, vp, vp, vp, i32]),
        "uaamg_poisson_apply": (C.c_int, [vp, C.c_int, C.c_int, f32, vp, vp, vp]),
        "uaamg_poisson_first_sweep": (C.c_int, [vp, C.c_int, C.c_int, f32, vp, vp]),
        "uaamg_poisson_restrict": (C.c_int, [vp, C.c_int, vp, vp]),
        "uaamg_poisson_prolong": (C.c_int, [vp, C.c_int, vp, vp, f32]),
        "uaamg_poisson_precondition": (C.c_int, [vp, vp, vp]),
        "uaamg_poisson_coarsest_solve": (C.c_int, [vp, vp, vp]),
        "uaamg_poisson_algorithmic_bytes": (C.c_int, [vp, C.c_int, C.c_int, P(C.c_double)]),
        "uaamg_poisson_time_apply": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(f32)]),
        "uaamg_scene_create": (C.c_int, [C.c_int, i32, C.c_uint64, P(vp)]),
        "uaamg_scene_create_custom": (C.c_int, [i32, f32, i32, vp, i32, vp, C.c_uint64, P(vp)]),
        "uaamg_scene_grids": (C.c_int, [vp, P(_Grid), P(_Grid), P(_Grid), P(_Grid), P(_Grid), P(f32), P(i32)]),
        "uaamg_scene_destroy": (None, [vp]),
        "uaamg_fraction_inside2": (f32, [f32, f32]),
        "uaamg_fraction_inside4": (f32, [f32, f32, f32, f32]),
        "uaamg_viscosity_options_default": (C.c_int, [P(_ViscOptions)]),
        "uaamg_viscosity_create": (C.c_int, [P(_Grid), P(_Grid), P(_Grid), f32, f32, f32, P(_ViscOptions), P(vp)]),
        "uaamg_viscosity_destroy": (None, [vp]),
        "uaamg_viscosity_set_options": (C.c_int, [vp, P(_ViscOptions)]),
        "uaamg_viscosity_ndof": (i64, [vp]),
        "uaamg_viscosity_nlevels": (C.c_int, [vp]),
        "uaamg_viscosity_level_info": (C.c_int, [vp, C.c_int, P(i64), vp, P(i64)]),
        "uaamg_viscosity_level_topology": (C.c_int, [vp, C.c_int, vp, vp, vp]),
        "uaamg_viscosity_level_coefficients": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
        "uaamg_viscosity_level_matrix": (C.c_int, [vp, C.c_int, P(i64), vp, vp, vp]),
        "uaamg_viscosity_build_rhs": (C.c_int, [vp, P(_Grid), P(_Grid), vp]),
        "uaamg_viscosity_build_rhs_packed": (C.c_int, [vp, P(_Grid), P(_Grid), P(_Grid), P(_Grid), vp]),
        "uaamg_viscosity_grid_to_vector": (C.c_int, [vp, P(_Grid), P(_Grid), P(_Grid), vp]),
        "uaamg_viscosity_vector_to_grid": (C.c_int, [vp, vp, vp]),
        "uaamg_packed3_from_vec3": (C.c_int, [P(_Grid), vp]),
        "uaamg_packed3_to_vec3": (C.c_int, [P(_Grid), P(_Grid), P(_Grid), P(i64), vp, vp, vp]),
        "uaamg_viscosity_solve_pcg": (C.c_int, [vp, f32, vp, vp, P(i32)]),
        "uaamg_viscosity_solve_pcg_device": (C.c_int, [vp, f32, vp, vp, P(i32)]),
        "uaamg_viscosity_apply": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp]),
        "uaamg_viscosity_restrict": (C.c_int, [vp, C.c_int, vp, vp]),
        "uaamg_viscosity_prolong": (C.c_int, [vp, C.c_int, vp, vp]),
        "uaamg_viscosity_coarsest_solve": (C.c_int, [vp, vp, vp]),
        "uaamg_viscosity_precondition": (C.c_int, [vp, vp, vp]),
        "uaamg_viscosity_algorithmic_bytes": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(C.c_double)]),
        "uaamg_viscosity_time_apply": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, P(f32)]),
        "uaamg_viscosity_kernel_map": (C.c_int, [vp, C.c_int, P(i32), P(f32), P(f32)]),
        "uaamg_viscosity_batch_create": (C.c_int, [i32, vp, vp, vp, vp, vp, vp, P(_ViscOptions), P(vp)]),
        "uaamg_viscosity_batch_destroy": (None, [vp]),
        "uaamg_viscosity_batch_info": (C.c_int, [vp, P(i32), P(i32), P(i64), P(i64)]),
        "uaamg_viscosity_batch_ndof": (i64, [vp, i32]),
        "uaamg_viscosity_batch_solve_device": (C.c_int, [vp, f32, vp, vp, vp]),
        "uaamg_visc_scene_create": (C.c_int, [C.c_int, i32, C.c_uint64, f32, f32, f32, P(vp)]),
        "uaamg_visc_scene_grids": (C.c_int, [vp, P(_Grid), P(_Grid), P(_Grid), P(_Grid), P(f32), P(f32), P(f32)]),
        "uaamg_visc_scene_destroy": (None, [vp]),
        "uaamg_volume_fraction8": (f32, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib_handle = L
    return L


def _check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().uaamg_last_error().decode(errors="replace")
        raise UaamgError(f"{what} failed ({rc}): {msg}")
    return rc


def device_count() -> int:
    return int(lib().uaamg_device_count())


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class Grid:
    """A flattened OpenVDB FloatGrid (ncomp=1) or Vec3fGrid (ncomp=3).

    origins: int32 [n,3]; active: uint64 [n,8] (bit b of word w = offset 64w+b);
    values: float32 [n,512] or [n,512,3] at voxel offset 64x+8y+z.
    """

    origins: np.ndarray
    values: np.ndarray
    active: np.ndarray | None = None
    background: tuple = (0.0, 0.0, 0.0)

    @property
    def ncomp(self) -> int:
        return 1 if self.values.ndim == 2 else 3

    def _struct(self, values_ptr: int | None = None) -> _Grid:
        """values_ptr: address of the values in device memory (same layout) instead of self.values"""
        o = np.ascontiguousarray(self.origins, dtype=np.int32)
        v = None if values_ptr is not None else np.ascontiguousarray(self.values, dtype=np.float32)
        a = None if self.active is None else np.ascontiguousarray(self.active, dtype=np.uint64)
        self._keep = (o, v, a)   # keep alive while the struct is in use
        g = _Grid()
        g.nleaf = o.shape[0]
        g.origins = o.ctypes.data
        g.values = values_ptr if values_ptr is not None else v.ctypes.data
        g.active = None if a is None else a.ctypes.data
        g.ncomp = self.ncomp
        bg = self.background if isinstance(self.background, (tuple, list)) else (self.background,) * 3
        for i in range(3):
            g.background[i] = float(bg[min(i, len(bg) - 1)])
        return g


def calculate_face_weights(liquid_sdf: Grid, solid_sdf: Grid, device: int = 0, *, member: bool = False) -> Grid:
    """FLIP_Solver_OpenVDB::calculate_face_weights (FLIP_vdb.cpp:1990-2068) with the weights computed on
    HIP device `device`: a Vec3f grid on the liquid SDF's topology dilated twice (26 neighbours), leaves
    in tree order, background (1, 1, 1). member=True: the member variant FLIP's substep calls
    (FLIP_vdb.cpp:2669-2752), which sets every weight below 1e-2 to 0."""
    gl, gs = liquid_sdf._struct(), solid_sdf._struct()
    variant = 1 if member else 0
    n = C.c_int64()
    _check(lib().uaamg_face_weights_ex(C.byref(gl), C.byref(gs), int(device), variant, C.byref(n), None, None, None),
           "face_weights")
    org = np.zeros((n.value, 3), np.int32)
    act = np.zeros((n.value, 8), np.uint64)
    val = np.zeros((n.value, 512, 3), np.float32)
    _check(lib().uaamg_face_weights_ex(C.byref(gl), C.byref(gs), int(device), variant, C.byref(n), _ptr(org),
                                       _ptr(act), _ptr(val)), "face_weights")
    return Grid(org, val, act, (1.0, 1.0, 1.0))


class Scene:
    """Synthetic scene (liquid SDF, face weights, face velocities) built by the library."""

    def __init__(self, config: int = 0, n: int = 0, seed: int = 0, *, custom: dict | None = None):
        L = lib()
        h = C.c_void_p()
        if custom is None:
            _check(L.uaamg_scene_create(int(config), int(n), C.c_uint64(seed), C.byref(h)), "scene_create")
        else:
            sph = np.ascontiguousarray(custom.get("spheres", np.zeros((0, 4))), dtype=np.float32).reshape(-1, 4)
            solid = custom.get("solid")
            sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.float32)
            _check(L.uaamg_scene_create_custom(int(custom["n"]), float(custom.get("pool", 0.0)), sph.shape[0],
                                               _ptr(sph), int(solid is not None), _ptr(sol),
                                               C.c_uint64(custom.get("seed", seed)), C.byref(h)), "scene_create")
        self._h = h
        gs = [_Grid() for _ in range(5)]
        dx = C.c_float()
        nn = C.c_int32()
        _check(L.uaamg_scene_grids(h, *[C.byref(g) for g in gs], C.byref(dx), C.byref(nn)), "scene_grids")
        self.dx = float(dx.value)
        self.n = int(nn.value)
        self.phi, self.weights, self.vx, self.vy, self.vz = (self._wrap(g) for g in gs)

    @staticmethod
    def _wrap(g: _Grid) -> Grid:
        n = g.nleaf
        org = np.ctypeslib.as_array(C.cast(g.origins, C.POINTER(C.c_int32)), shape=(max(n, 1) * 3,))[: n * 3]
        act = np.ctypeslib.as_array(C.cast(g.active, C.POINTER(C.c_uint64)), shape=(max(n, 1) * 8,))[: n * 8]
        shape = (n, 512) if g.ncomp == 1 else (n, 512, 3)
        size = max(n, 1) * 512 * g.ncomp
        val = np.ctypeslib.as_array(C.cast(g.values, C.POINTER(C.c_float)), shape=(size,))[: n * 512 * g.ncomp]
        return Grid(org.reshape(n, 3).copy(), val.reshape(shape).copy(), act.reshape(n, 8).copy(),
                    tuple(g.background))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().uaamg_scene_destroy(h)
            self._h = None


def solve_batch_device(laplacians, x_ptrs, b_ptrs, concurrency: int = 4):
    """Solves many problems at once (one LaplacianWithLevel each, device vectors given as raw
    pointers); returns (status list, iteration list)."""
    n = len(laplacians)
    hs = (C.c_void_p * n)(*[lap._h.value for lap in laplacians])
    xs = (C.c_void_p * n)(*x_ptrs)
    bs = (C.c_void_p * n)(*b_ptrs)
    st = np.zeros(n, np.int32)
    it = np.zeros(n, np.int32)
    _check(lib().uaamg_poisson_solve_batch_device(hs, n, xs, bs, _ptr(st), _ptr(it), int(concurrency)),
           "solve_batch_device")
    return st.tolist(), it.tolist()


class PoissonBatch:
    """Many Poisson problems merged into one leaf table per level and solved
    together (uaamg_poisson_batch_*): one launch per W-cycle stage for the
    whole batch, per-problem PCG scalars and convergence on the device.
    `problems` is a list of (liquid_phi, face_weights, dt, dx)."""

    def __init__(self, problems, *, precision: int = 0, flags: int = 0, smoother: int = 0,
                 tolerance: float = 1e-7, max_iterations: int = 100, device: int = 0):
        L = lib()
        n = len(problems)
        o = _Options()
        L.uaamg_options_default(C.byref(o))
        o.precision, o.flags, o.smoother, o.device = precision, flags, smoother, device
        o.relative_tolerance, o.max_iterations = float(tolerance), int(max_iterations)
        self._grids = [(p[0]._struct(), p[1]._struct()) for p in problems]
        phis = (C.c_void_p * n)(*[C.addressof(g[0]) for g in self._grids])
        ws = (C.c_void_p * n)(*[C.addressof(g[1]) for g in self._grids])
        dts = np.array([p[2] for p in problems], np.float32)
        dxs = np.array([p[3] for p in problems], np.float32)
        h = C.c_void_p()
        _check(L.uaamg_poisson_batch_create(n, phis, ws, _ptr(dts), _ptr(dxs), C.byref(o), C.byref(h)),
               "uaamg_poisson_batch_create")
        self._h = h
        self.n = n

    def set_options(self, tolerance: float, max_iterations: int = 100):
        o = _Options()
        lib().uaamg_options_default(C.byref(o))
        o.relative_tolerance, o.max_iterations = float(tolerance), int(max_iterations)
        _check(lib().uaamg_poisson_batch_set_options(self._h, C.byref(o)), "batch set_options")

    def info(self) -> dict:
        n, nl = C.c_int32(), C.c_int32()
        l0, nb = C.c_int64(), C.c_int64()
        _check(lib().uaamg_poisson_batch_info(self._h, C.byref(n), C.byref(nl), C.byref(l0), C.byref(nb)), "info")
        return {"nproblems": n.value, "nlevels": nl.value, "leaves0": l0.value, "device_bytes": nb.value}

    def ndof(self, i: int) -> int:
        return int(lib().uaamg_poisson_batch_ndof(self._h, i))

    def solve_device(self, x_ptrs, b_ptrs):
        """x_ptrs / b_ptrs: device addresses of each problem's DOF vectors; returns (status, iterations)."""
        xs = (C.c_void_p * self.n)(*x_ptrs)
        bs = (C.c_void_p * self.n)(*b_ptrs)
        st = np.zeros(self.n, np.int32)
        it = np.zeros(self.n, np.int32)
        _check(lib().uaamg_poisson_batch_solve_device(self._h, xs, bs, _ptr(st), _ptr(it)), "batch solve")
        return st.tolist(), it.tolist()

    def last_residual(self):
        """Per problem: (FP64 true residual, residual replacements) of the last solve; -1s without replacement."""
        r = np.zeros(self.n, np.float32)
        k = np.zeros(self.n, np.int32)
        _check(lib().uaamg_poisson_batch_last_residual(self._h, _ptr(r), _ptr(k)), "batch last_residual")
        return list(zip(r.tolist(), k.tolist()))

    def last_fallback(self):
        """Per problem: (status, iterations) of the pure-multigrid fallback of the last solve (flags
        bit 7, FLIP_vdb.cpp:3150-3155); (-1, -1) where the PCG succeeded."""
        st = np.zeros(self.n, np.int32)
        it = np.zeros(self.n, np.int32)
        _check(lib().uaamg_poisson_batch_last_fallback(self._h, _ptr(st), _ptr(it)), "batch last_fallback")
        return list(zip(st.tolist(), it.tolist()))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().uaamg_poisson_batch_destroy(h)
            self._h = None


PRECISION_FP32 = 0         # FP32 throughout (the reference's arithmetic)
PRECISION_BF16_CYCLE = 1   # BF16-stored W-cycle vectors, FP32 arithmetic, FP64 residual replacement


class SmootherOption:
    """PoissonSolver::SmootherOption (simd_vdb_poisson.h:143-148)."""

    ScheduledRelaxedJacobi = 0
    WeightedJacobi = 1
    RedBlackGaussSeidel = 2
    SPAI0 = 3


class LaplacianWithLevel:
    """Level-0 pressure Laplacian plus its aggregation hierarchy on the GPU.

    createPressurePoissonLaplacian(liquid_phi, face_weights, dt) mirrors
    simd_vdb_poisson.h:54-57; dx is liquid_phi's voxel size.
    """

    def __init__(self, liquid_phi: Grid, face_weights: Grid, dt: float, dx: float, *, device: int = 0,
                 max_coarsest_dof: int = 4000, flags: int = 0, precision: int = 0, smoother: int | None = None):
        L = lib()
        o = _Options()
        L.uaamg_options_default(C.byref(o))
        if smoother is not None:
            o.smoother = int(smoother)   # the smoother the handle is created with (PoissonSolver may change it)
        o.device = device
        o.max_coarsest_dof = max_coarsest_dof
        o.flags = flags
        o.precision = precision
        h = C.c_void_p()
        gphi, gw = liquid_phi._struct(), face_weights._struct()
        _check(L.uaamg_poisson_create(C.byref(gphi), C.byref(gw), float(dt), float(dx), C.byref(o), C.byref(h)),
               "uaamg_poisson_create")
        self._h = h
        self._device = device
        self.precision = precision
        self.mDt = float(dt)
        self.mDxThisLevel = float(dx)
        self.mLevel = 0
        self.mNumDof = int(L.uaamg_poisson_ndof(h))

    @classmethod
    def createPressurePoissonLaplacian(cls, liquid_phi: Grid, face_weights: Grid, dt: float, dx: float | None = None,
                                       **kw) -> "LaplacianWithLevel":
        if dx is None:
            raise ValueError("dx (the liquid phi voxel size) is required")
        return cls(liquid_phi, face_weights, dt, dx, **kw)

    def createPressurePoissonRightHandSide(self, face_weights: Grid, vx: Grid, vy: Grid, vz: Grid,
                                           solid_velocity: Grid | None = None, dt: float | None = None) -> np.ndarray:
        rhs = np.zeros(self.mNumDof, dtype=np.float32)
        g = [face_weights._struct(), vx._struct(), vy._struct(), vz._struct()]
        sv = None if solid_velocity is None else C.byref(solid_velocity._struct())
        _check(lib().uaamg_poisson_build_rhs(self._h, *[C.byref(x) for x in g], sv, _ptr(rhs)), "build_rhs")
        return rhs

    def createPressurePoissonRightHandSide_device(self, rhs_ptr: int, face_weights: Grid, vx: Grid, vy: Grid,
                                                  vz: Grid, solid_velocity: Grid | None = None,
                                                  values_ptrs: list | None = None) -> None:
        """The right-hand side built on the device into rhs_ptr (ndof floats, DOF order, device memory).
        values_ptrs: device addresses of the grids' values, in argument order (None: upload
        each grid's host values for the call)."""
        grids = [face_weights, vx, vy, vz] + ([] if solid_velocity is None else [solid_velocity])
        if values_ptrs is not None and len(values_ptrs) != len(grids):
            raise ValueError("one device values pointer per grid")
        st = [g._struct(None if values_ptrs is None else values_ptrs[i]) for i, g in enumerate(grids)]
        sv = None if solid_velocity is None else C.byref(st[4])
        _check(lib().uaamg_poisson_build_rhs_device(self._h, *[C.byref(x) for x in st[:4]], sv,
                                                    0 if values_ptrs is None else 1, C.c_void_p(rhs_ptr)),
               "build_rhs_device")

    def apply_pressure_gradient_device(self, pressure_ptr: int, liquid_phi: Grid, face_weights: Grid,
                                       solid_velocity: Grid | None, velocity: list, dt: float, friction: float,
                                       values_on_device: bool = False):
        """FLIP_Solver_OpenVDB::apply_pressure_gradient (FLIP_vdb.cpp:2571-2667) on the device, the
        pressure being the solve's DOF-ordered result in device memory (pressure_ptr). velocity: three
        Grids (x, y, z faces; values [n,512], active [n,8]). Returns per channel (active, values,
        update, reflected) after the update, in the channel's leaf order; the inputs are not modified.
        values_on_device: stage every values array in device memory (torch) and let the library work
        on it there (the device-resident calling form)."""
        keep = []

        def place(a: np.ndarray) -> int:
            if not values_on_device:
                return a.ctypes.data
            import torch
            t = torch.from_numpy(a).to(torch.device("cuda", self._device))
            keep.append((a, t))
            return t.data_ptr()

        grids = [liquid_phi, face_weights] + ([] if solid_velocity is None else [solid_velocity])
        st = [g._struct(place(np.ascontiguousarray(g.values, dtype=np.float32)) if values_on_device else None)
              for g in grids]
        sv = None if solid_velocity is None else C.byref(st[2])
        chans = (_FaceChannel * 3)()
        out = []
        for c, g in enumerate(velocity):
            org = np.ascontiguousarray(g.origins, dtype=np.int32)
            n = org.shape[0]
            act = (np.full((n, 8), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64) if g.active is None
                   else np.array(g.active, dtype=np.uint64, copy=True))
            val = np.array(g.values, dtype=np.float32, copy=True).reshape(n, 512)
            upd = np.zeros_like(val)
            refl = np.zeros_like(val)
            chans[c].nleaf = n
            chans[c].origins = org.ctypes.data
            chans[c].active = act.ctypes.data
            chans[c].values = place(val)
            chans[c].update = place(upd)
            chans[c].reflected = place(refl)
            out.append((org, act, val, upd, refl))
        _check(lib().uaamg_poisson_apply_pressure_gradient_device(self._h, C.c_void_p(pressure_ptr), C.byref(st[0]),
                                                                  C.byref(st[1]), sv, chans, float(dt),
                                                                  float(friction), 1 if values_on_device else 0),
               "apply_pressure_gradient_device")
        if values_on_device:
            import torch
            torch.cuda.synchronize()
            for a, t in keep:
                a[...] = t.cpu().numpy()   # values / update / reflected come back; inputs are copies
        return [(a, v, u, r) for (_, a, v, u, r) in out]

    def getZeroVectorGrid(self) -> np.ndarray:
        return np.zeros(self.mNumDof, dtype=np.float32)

    # ---- hierarchy introspection
    @property
    def nlevels(self) -> int:
        return int(lib().uaamg_poisson_nlevels(self._h))

    def level_info(self, level: int) -> dict:
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().uaamg_poisson_level_info(self._h, level, C.byref(a), C.byref(b), C.byref(c)), "level_info")
        return {"ndof": a.value, "nleaf_dof": b.value, "nleaf_tree": c.value}

    def level_topology(self, level: int):
        n = self.level_info(level)["nleaf_dof"]
        org = np.zeros((n, 3), np.int32)
        mask = np.zeros((n, 8), np.uint64)
        base = np.zeros(n, np.int64)
        _check(lib().uaamg_poisson_level_topology(self._h, level, _ptr(org), _ptr(mask), _ptr(base)), "topology")
        return org, mask, base

    def level_coefficients(self, level: int, kind: int):
        n = self.level_info(level)["nleaf_dof"]
        br = np.zeros((n, 512), np.float32)
        kept = np.zeros(n, np.uint8)
        _check(lib().uaamg_poisson_level_coefficients(self._h, level, kind, _ptr(br), _ptr(kept)), "coefficients")
        return br, kept

    def level_ndof(self, level: int) -> int:
        return self.level_info(level)["ndof"]

    # ---- single operators on DOF-ordered vectors (LaplacianWithLevel methods)
    def apply(self, level: int, mode: int, x: np.ndarray, b: np.ndarray | None = None, weight: float = 0.0):
        x = np.ascontiguousarray(x, np.float32)
        b = None if b is None else np.ascontiguousarray(b, np.float32)
        y = np.zeros(self.level_ndof(level), np.float32)
        _check(lib().uaamg_poisson_apply(self._h, level, mode, float(weight), _ptr(x), _ptr(b), _ptr(y)), "apply")
        return y

    def laplacianApply(self, x, level: int = 0):
        return self.apply(level, 0, x)

    def residualApply(self, x, b, level: int = 0):
        return self.apply(level, 1, x, b)

    def weightedJacobiApply(self, x, b, weight: float = 6.0 / 7.0, level: int = 0):
        return self.apply(level, 2, x, b, weight)

    def spai0Apply(self, x, b, level: int = 0):
        return self.apply(level, 3, x, b)

    def first_sweep(self, level: int, kind: int, b: np.ndarray, weight: float) -> np.ndarray:
        b = np.ascontiguousarray(b, np.float32)
        y = np.zeros(self.level_ndof(level), np.float32)
        _check(lib().uaamg_poisson_first_sweep(self._h, level, kind, float(weight), _ptr(b), _ptr(y)), "first")
        return y

    def restriction(self, fine_level: int, fine: np.ndarray) -> np.ndarray:
        fine = np.ascontiguousarray(fine, np.float32)
        out = np.zeros(self.level_ndof(fine_level + 1), np.float32)
        _check(lib().uaamg_poisson_restrict(self._h, fine_level, _ptr(fine), _ptr(out)), "restrict")
        return out

    def prolongation(self, fine_level: int, coarse: np.ndarray, fine: np.ndarray, alpha: float = 1.0):
        coarse = np.ascontiguousarray(coarse, np.float32)
        out = np.array(fine, dtype=np.float32, copy=True)
        _check(lib().uaamg_poisson_prolong(self._h, fine_level, _ptr(coarse), _ptr(out), float(alpha)), "prolong")
        return out

    def algorithmic_bytes(self, level: int, mode: int) -> float:
        v = C.c_double()
        _check(lib().uaamg_poisson_algorithmic_bytes(self._h, level, mode, C.byref(v)), "algorithmic_bytes")
        return float(v.value)

    def time_apply(self, level: int, mode: int, reps: int = 50) -> float:
        """Mean ms per launch of one operator sweep, HIP events on the solver stream."""
        v = C.c_float()
        _check(lib().uaamg_poisson_time_apply(self._h, level, mode, reps, C.byref(v)), "time_apply")
        return float(v.value)

    def solve_pcg_device(self, x_ptr: int, b_ptr: int) -> tuple[int, int]:
        """solveMultigridPCG on DOF-ordered device vectors given as raw device pointers."""
        it = C.c_int32()
        rc = _check(lib().uaamg_poisson_solve_pcg_device(self._h, C.c_void_p(x_ptr), C.c_void_p(b_ptr),
                                                         C.byref(it)), "solve_pcg_device")
        return rc, int(it.value)

    def getTriplets(self, level: int = 0):
        """(rows, cols, vals) of the level's explicit matrix in getTriplets order."""
        n = C.c_int64()
        _check(lib().uaamg_poisson_level_triplets(self._h, level, C.byref(n), None, None, None), "triplets")
        r = np.zeros(n.value, np.int32)
        c = np.zeros(n.value, np.int32)
        v = np.zeros(n.value, np.float32)
        _check(lib().uaamg_poisson_level_triplets(self._h, level, C.byref(n), _ptr(r), _ptr(c), _ptr(v)), "triplets")
        return r, c, v

    def kernel_map(self) -> list[dict]:
        """Per level: whether visits run as block-visit launches (post_cg: the post-visit launch also
        runs the coarsest solve), and the timings behind the choice."""
        out = []
        for l in range(self.nlevels):
            b, mb, ms = C.c_int32(), C.c_float(), C.c_float()
            _check(lib().uaamg_poisson_kernel_map(self._h, l, C.byref(b), C.byref(mb), C.byref(ms)), "kernel_map")
            out.append({"level": l, "block": bool(b.value), "post_cg": b.value == 2, "ms_block": round(mb.value, 4),
                        "ms_sweep": round(ms.value, 4)})
        return out

    def last_residual(self) -> tuple[float, int]:
        """(FP64 ||b - A x||_inf of the last solve, residual replacements made); -1s without replacement."""
        r, k = C.c_float(), C.c_int32()
        _check(lib().uaamg_poisson_last_residual(self._h, C.byref(r), C.byref(k)), "last_residual")
        return float(r.value), int(k.value)

    def coarsest_solve(self, b: np.ndarray) -> np.ndarray:
        b = np.ascontiguousarray(b, np.float32)
        x = np.zeros(self.level_ndof(self.nlevels - 1), np.float32)
        _check(lib().uaamg_poisson_coarsest_solve(self._h, _ptr(b), _ptr(x)), "coarsest")
        return x

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().uaamg_poisson_destroy(h)
            self._h = None


class PoissonSolver:
    """simd_uaamg::PoissonSolver over a LaplacianWithLevel (simd_vdb_poisson.h:139-205)."""

    SUCCESS = SUCCESS
    FAILED = FAILED

    def __init__(self, lhs: LaplacianWithLevel):
        self.mMultigridLevel0 = lhs
        self.mIterationTaken = 0
        self.mMaxIteration = 100
        self.mRelativeTolerance = 1e-7
        self.mSmoother = SmootherOption.ScheduledRelaxedJacobi

    def _push_options(self):
        o = _Options()
        lib().uaamg_options_default(C.byref(o))
        o.relative_tolerance = float(self.mRelativeTolerance)
        o.max_iterations = int(self.mMaxIteration)
        o.smoother = int(self.mSmoother)
        _check(lib().uaamg_poisson_set_options(self.mMultigridLevel0._h, C.byref(o)), "set_options")

    def solveMultigridPCG(self, in_out_pressure: np.ndarray, rhs: np.ndarray) -> int:
        """In-place on in_out_pressure (float32, length mNumDof); returns SUCCESS/FAILED."""
        self._push_options()
        if in_out_pressure.dtype != np.float32 or not in_out_pressure.flags.c_contiguous:
            raise TypeError("pressure must be a contiguous float32 array")
        b = np.ascontiguousarray(rhs, np.float32)
        it = C.c_int32()
        rc = _check(lib().uaamg_poisson_solve_pcg(self.mMultigridLevel0._h, _ptr(in_out_pressure), _ptr(b),
                                                  C.byref(it)), "solveMultigridPCG")
        self.mIterationTaken = int(it.value)
        return rc

    def solvePureMultigrid(self, in_out_pressure: np.ndarray, rhs: np.ndarray) -> int:
        self._push_options()
        if in_out_pressure.dtype != np.float32 or not in_out_pressure.flags.c_contiguous:
            raise TypeError("pressure must be a contiguous float32 array")
        b = np.ascontiguousarray(rhs, np.float32)
        it = C.c_int32()
        rc = _check(lib().uaamg_poisson_solve_mg(self.mMultigridLevel0._h, _ptr(in_out_pressure), _ptr(b),
                                                 C.byref(it)), "solvePureMultigrid")
        self.mIterationTaken = int(it.value)
        return rc

    def precondition(self, r: np.ndarray) -> np.ndarray:
        """One W-cycle z = M r (muCyclePreconditioner<2,true>)."""
        self._push_options()
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros_like(r)
        _check(lib().uaamg_poisson_precondition(self.mMultigridLevel0._h, _ptr(r), _ptr(z)), "precondition")
        return z


from .viscosity import L_with_level, ViscBatch, ViscScene, packed_FloatGrid3, simd_viscosity3d  # noqa: E402
