"""Python binding of the solve server (uaamg_server_* in include/uaamg_capi.h).

The server is a C++ request loop over many resident problems on one GPU:
problems are registered once, requests (a right-hand side in device memory)
are queued, and worker threads admit them as slots and device memory free up,
making the problem resident (hierarchy, PCG vectors, warm-start solution) and
evicting idle problems least recently used first when the memory budget
requires it. Solves of different problems overlap on the device.
"""
from __future__ import annotations

import ctypes as C

from . import Grid, UaamgError, _Options, _ViscOptions, _check, lib

SUCCESS, FAILED = 0, 1


class _ServerOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("memory_budget_bytes", C.c_int64)]


class _ServerCounters(C.Structure):
    _fields_ = [(k, C.c_int64) for k in
                ("completed", "failed", "evictions", "acquisitions", "resident_bytes", "peak_inflight", "queued")]


_SIGS = False


def _bind():
    global _SIGS
    if _SIGS:
        return lib()
    L = lib()
    vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
    P = C.POINTER
    sig = {
        "uaamg_server_options_default": (C.c_int, [P(_ServerOptions)]),
        "uaamg_server_create": (C.c_int, [P(_ServerOptions), P(vp)]),
        "uaamg_server_destroy": (None, [vp]),
        "uaamg_server_add_poisson": (C.c_int, [vp, vp, vp, f32, f32, P(_Options), P(i64)]),
        "uaamg_server_add_viscosity": (C.c_int, [vp, vp, vp, vp, f32, f32, f32, P(_ViscOptions), P(i64)]),
        "uaamg_server_problem_info": (C.c_int, [vp, i64, P(i64), P(i64), P(i32)]),
        "uaamg_server_submit": (C.c_int, [vp, i64, vp, vp, vp, f32, P(i64)]),
        "uaamg_server_wait": (C.c_int, [vp, i64, P(i32), P(i32), P(C.c_double), P(C.c_double)]),
        "uaamg_server_wait_fallback": (C.c_int, [vp, i64, P(i32), P(i32), P(i32), P(i32), P(C.c_double),
                                                 P(C.c_double)]),
        "uaamg_server_stats": (C.c_int, [vp, P(_ServerCounters)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _SIGS = True
    return L


class SolveServer:
    """A solve server on one device with `slots` concurrent solves and a memory
    budget in bytes (0: 90% of the free device memory)."""

    def __init__(self, device: int = 0, slots: int = 4, memory_budget: int = 0):
        L = _bind()
        o = _ServerOptions()
        L.uaamg_server_options_default(C.byref(o))
        o.device, o.slots, o.memory_budget_bytes = device, slots, int(memory_budget)
        h = C.c_void_p()
        _check(L.uaamg_server_create(C.byref(o), C.byref(h)), "uaamg_server_create")
        self._h = h

    def add_poisson(self, liquid_phi: Grid, face_weights: Grid, dt: float, dx: float, *, precision: int = 0,
                    flags: int = 0, smoother: int = 0, max_iterations: int = 100) -> int:
        L = _bind()
        o = _Options()
        L.uaamg_options_default(C.byref(o))
        o.precision, o.flags, o.smoother, o.max_iterations = precision, flags, smoother, max_iterations
        pid = C.c_int64()
        gp, gw = liquid_phi._struct(), face_weights._struct()
        _check(L.uaamg_server_add_poisson(self._h, C.byref(gp), C.byref(gw), float(dt), float(dx), C.byref(o),
                                          C.byref(pid)), "uaamg_server_add_poisson")
        return int(pid.value)

    def add_viscosity(self, viscosity: Grid, liquid_sdf: Grid, solid_sdf: Grid, dt: float, rho: float, dx: float, *,
                      flags: int = 0, max_iterations: int = 100, precision: int = 0) -> int:
        L = _bind()
        o = _ViscOptions()
        L.uaamg_viscosity_options_default(C.byref(o))
        o.flags, o.max_iterations, o.precision = flags, max_iterations, precision
        pid = C.c_int64()
        g = [viscosity._struct(), liquid_sdf._struct(), solid_sdf._struct()]
        _check(L.uaamg_server_add_viscosity(self._h, *[C.byref(x) for x in g], float(dt), float(rho), float(dx),
                                            C.byref(o), C.byref(pid)), "uaamg_server_add_viscosity")
        return int(pid.value)

    def problem_info(self, pid: int) -> dict:
        nd, nb, res = C.c_int64(), C.c_int64(), C.c_int32()
        _check(_bind().uaamg_server_problem_info(self._h, pid, C.byref(nd), C.byref(nb), C.byref(res)),
               "problem_info")
        return {"ndof": nd.value, "device_bytes": nb.value, "resident": bool(res.value)}

    def submit(self, pid: int, b_ptr: int, x_out_ptr: int = 0, x0_ptr: int = 0, tol: float = 1e-7) -> int:
        """Queues a solve; pointers are device addresses (0 = none) that stay valid until wait()."""
        t = C.c_int64()
        _check(_bind().uaamg_server_submit(self._h, pid, C.c_void_p(b_ptr), C.c_void_p(x_out_ptr or None),
                                           C.c_void_p(x0_ptr or None), float(tol), C.byref(t)), "submit")
        return int(t.value)

    def wait(self, ticket: int) -> tuple[int, int, float, float]:
        """(status, iterations, queue ms, total ms) of a ticket; blocks until it completes."""
        st, it = C.c_int32(), C.c_int32()
        q, tot = C.c_double(), C.c_double()
        _check(_bind().uaamg_server_wait(self._h, ticket, C.byref(st), C.byref(it), C.byref(q), C.byref(tot)),
               "wait")
        return int(st.value), int(it.value), float(q.value), float(tot.value)

    @staticmethod
    def last_error() -> str:
        """After wait() returned a negative status on this thread: the failed step and its HIP error."""
        return _bind().uaamg_last_error().decode(errors="replace")

    def wait_fallback(self, ticket: int) -> tuple[int, int, int, int, float, float]:
        """(status, iterations, fallback status, fallback iterations, queue ms, total ms): wait() plus
        the pure-multigrid fallback of a problem registered with flags bit 7 (-1s when it did not run)."""
        st, it, ms, mi = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        q, tot = C.c_double(), C.c_double()
        _check(_bind().uaamg_server_wait_fallback(self._h, ticket, C.byref(st), C.byref(it), C.byref(ms),
                                                  C.byref(mi), C.byref(q), C.byref(tot)), "wait_fallback")
        return int(st.value), int(it.value), int(ms.value), int(mi.value), float(q.value), float(tot.value)

    def stats(self) -> dict:
        c = _ServerCounters()
        _check(_bind().uaamg_server_stats(self._h, C.byref(c)), "stats")
        return {k: int(getattr(c, k)) for k, _ in _ServerCounters._fields_}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _bind().uaamg_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


__all__ = ["SolveServer", "UaamgError", "SUCCESS", "FAILED"]
