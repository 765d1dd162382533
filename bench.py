#!/usr/bin/env python3
"""Benchmark of the UAAMG-PCG pressure-Poisson solve on one MI355X.

Workload (default): the 256^3 free-surface scene of the headline metric - a 30%
pool plus 8 seeded liquid spheres (r in [0.05, 0.2]) around a static solid
sphere (r = 0.2 at the centre) that gives fractional face weights; rhs = the
divergence of seeded uniform[-1,1] face velocities; zero initial guess;
solveMultigridPCG to 1e-7 relative residual (infinity norm) with the
scheduled-relaxed-Jacobi W-cycle. Three forms are timed: plain FP32 (the
reference's arithmetic), FP32 with the FP64 residual replacement, and the
measured form - BF16-stored W-cycle vectors with FP32 arithmetic, FP64 dots
and the FP64 residual replacement.

A step is one full solve with the hierarchy built and b already resident in
HBM; the timed region covers K solves (each zeroes x on the device first).
The JSON line also carries:
  roofline      the kernel that takes the most time per solve among the
                level-0 sweep, each coarse level's block post-visit kernel and
                the coarsest CG (launch time measured with HIP events on the
                solver's stream times the launches per solve the W-cycle
                makes): its algorithmic bytes per launch / mean launch time,
                against the 8 TB/s HBM peak;
  roofline_solve the whole solve: the algorithmic bytes of every operation of
                the solve (per-DOF stream counts, coefficients only where
                stored) / ms per solve;
  roofline_level0_sweep  the level-0 weighted Jacobi sweep;
  cpu_baseline  the CPU oracle (a C++ restatement of the reference's
                SIMD/TBB path) solving the same system on the host cores:
                median of 5 solves after one warm-up.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

METRIC = "ms/Poisson solve to 1e-7 rel. residual @256³ vs plain FP32; solves/s + p99 latency"
HBM_PEAK_GBS = 8000.0


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/hbm_traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled
    by the copy-kernel calibration), or None when it has not been measured."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    entry = table.get("kernels", {}).get(kernel)
    return None if entry is None else entry.get("bytes_per_launch")


def solve_bytes(lap, iterations: int, replacements: int, mixed: bool) -> float:
    """Algorithmic bytes of one PCG solve: every operation of the solve counted once per launch
    with the per-DOF stream bytes of each array it must read or write (coefficient streams only
    where the level stores them), unfused - the kernels' fusions only lower the real traffic.
    W-cycle vectors are 2 bytes in the mixed form (level 0's right-hand side, the PCG residual,
    stays 4); the outer PCG is FP32."""
    L = lap.nlevels
    n = [lap.level_info(l)["ndof"] for l in range(L)]
    cs = [lap.algorithmic_bytes(l, 0) - 8.0 * n[l] for l in range(L)]            # sum of c over DOFs
    ds = [lap.algorithmic_bytes(l, 2) - lap.algorithmic_bytes(l, 1) for l in range(L)]   # sum of d
    sx = [2.0 if mixed else 4.0 for l in range(L)]
    sb = [4.0 if l == 0 else (2.0 if mixed else 4.0) for l in range(L)]

    def jacobi(l):
        return n[l] * (2 * sx[l] + sb[l]) + cs[l] + ds[l]

    def visit(l, from_zero):
        pre = (n[l] * (sb[l] + sx[l]) + ds[l] + 2 * jacobi(l)) if from_zero else 3 * jacobi(l)
        res = n[l] * (2 * sx[l] + sb[l]) + cs[l]
        rst = n[l] * sx[l] + n[l + 1] * sb[l + 1]
        pro = 2 * n[l] * sx[l] + n[l + 1] * sx[l + 1]
        return pre + res + rst + pro + 3 * jacobi(l)

    wcycle = visit(0, True)
    for l in range(1, L - 1):
        wcycle += 2 ** (l - 1) * (visit(l, True) + visit(l, False))
    wcycle += 2 ** max(L - 2, 0) * n[L - 1] * (sb[L - 1] + sx[L - 1] + 60.0)   # coarsest: b, x, its rows
    n0, c0 = n[0], cs[0]
    init = n0 * 12 + c0 + n0 * 4                       # r = b - A x, ||r||_inf
    first = 8 * n0 + (8 * n0 + c0) + 8 * n0 + 12 * n0  # rho, A p, sigma, r -= alpha A p with ||r||_inf
    per_iter = 8 * n0 + 24 * n0 + (8 * n0 + c0) + 8 * n0 + 12 * n0
    replaced = replacements * (4 * n0 + c0)            # r = b - A (x + alpha p) in place of the axpy
    return init + (iterations + 1) * wcycle + first + iterations * per_iter + replaced + 12 * n0


def kernel_rooflines(lap, iterations: int, reps: int) -> list:
    """Live launch time (HIP events on the solver's stream) and algorithmic bytes of the kernels
    that can dominate a solve: the level-0 sweep, each block-visit level's post-visit kernel and
    the coarsest CG; with the launches per solve the W-cycle makes of each."""
    L = lap.nlevels
    W = iterations + 1
    km = lap.kernel_map()
    out = []
    mixed = lap.precision == 1
    mode0 = 6 if mixed else 2
    ms = lap.time_apply(0, mode0, reps=reps)
    out.append({"kernel": ("apply_kernel<2, true, bf16, float, bf16>" if mixed else "apply_kernel<2, true>") +
                          " (level-0 Jacobi sweep)", "bytes_per_launch": lap.algorithmic_bytes(0, mode0),
                "ms_per_launch": ms, "launches_per_solve": 2 * W})
    vt = "bf16" if mixed else "float"
    for l in range(1, L - 1):
        if not km[l]["block"]:
            continue
        ms = lap.time_apply(l, 7, reps=reps)
        name = (f"block_post_cg_kernel<{vt}> (level-{l} post-visit with the level-{L - 1} coarsest CG in front)"
                if km[l]["post_cg"] else
                f"block_visit_kernel<3, 2, {vt}> (level-{l} post-visit: prolongation + 3 sweeps)")
        out.append({"kernel": name, "bytes_per_launch": lap.algorithmic_bytes(l, 7),
                    "ms_per_launch": ms, "launches_per_solve": W * 2 ** l})
    # the standalone coarsest CG runs only when the level above does not take it into its post-visit
    ms = lap.time_apply(L - 1, 8, reps=reps)
    out.append({"kernel": f"coarsest_cg_kernel<{vt}> (level-{L - 1} Eigen-CG restatement)",
                "bytes_per_launch": lap.algorithmic_bytes(L - 1, 8), "ms_per_launch": ms,
                "launches_per_solve": 0 if L >= 3 and km[L - 2]["post_cg"] else W * 2 ** max(L - 2, 0)})
    for k in out:
        k["ms_per_solve"] = k["ms_per_launch"] * k["launches_per_solve"]
    return out


def roofline_entry(k: dict, traffic=None) -> dict:
    ach = k["bytes_per_launch"] / (k["ms_per_launch"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": k["kernel"],
            "bytes_per_launch": k["bytes_per_launch"], "ms_per_launch": round(k["ms_per_launch"], 5),
            "launches_per_solve": k.get("launches_per_solve"),
            "ms_per_solve_est": round(k.get("ms_per_solve", 0.0), 4)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--smoother", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--workload", choices=["poisson", "viscosity", "batch", "visc-batch", "serve"],
                    default="poisson",
                    help="poisson: the headline line (plus viscosity / batch / serving objects unless --no-*); "
                         "viscosity, batch, serve: only that measurement (profiling runs)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-precision", type=int, default=1, help="Poisson precision in the serving pool")
    ap.add_argument("--serve-budget-gb", type=float, default=0.0, help="server memory budget (0: 90%% of free)")
    ap.add_argument("--no-viscosity", action="store_true")
    ap.add_argument("--no-512", action="store_true", help="skip the 512^3 object")
    ap.add_argument("--profile-form", choices=["fp32", "bf16"], default=None,
                    help="only time the solves of one Poisson form (for a kernel trace)")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="skip the per-kernel timing launches (for a profiled run of the solves alone)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-size", type=int, default=64)
    ap.add_argument("--batch-concurrency", type=int, default=4)
    ap.add_argument("--visc-batch-size", type=int, default=64)
    ap.add_argument("--visc-n", type=int, default=128)
    ap.add_argument("--visc-steps", type=int, default=10)
    ap.add_argument("--kernel-only", action="store_true",
                    help="build the workload, then only run the level-0 Jacobi sweep --kernel-reps times and a "
                         "256 MiB device copy 10 times (short run for hardware-counter passes)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: the solver is a single-GPU path and multi-GPU runs are "
              "not part of this build (use --gpus 1).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import uaamg_amd as U

    if U.device_count() <= 0 or not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 3
    dev = torch.device("cuda:0")

    if args.workload == "serve":
        out = {"serving": run_serving(args, dev)}
        out["serving_tight_budget"] = run_tight_serving(args, dev, out["serving"])
        print(json.dumps(out), flush=True)
        return 0
    if args.workload != "poisson":
        fn = {"viscosity": run_viscosity, "batch": run_batch, "visc-batch": run_visc_batch}[args.workload]
        print(json.dumps(fn(args, dev)), flush=True)
        return 0

    t0 = time.perf_counter()
    scene = U.Scene(args.config, args.n, seed=args.seed)
    t_scene = time.perf_counter() - t0
    t0 = time.perf_counter()
    lap = U.LaplacianWithLevel(scene.phi, scene.weights, dt=1.0, dx=scene.dx, device=0)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    b_host = lap.createPressurePoissonRightHandSide(scene.weights, scene.vx, scene.vy, scene.vz)
    t_rhs_host = time.perf_counter() - t0

    def make_solver(l):
        sol = U.PoissonSolver(l)
        sol.mRelativeTolerance = args.tol
        sol.mSmoother = args.smoother
        sol._push_options()
        return sol

    solver = make_solver(lap)
    b = torch.from_numpy(b_host).to(dev)
    x = torch.zeros_like(b)
    iters = None
    if args.kernel_only:
        ms = lap.time_apply(0, 2, reps=args.kernel_reps)
        lap_k = U.LaplacianWithLevel(scene.phi, scene.weights, dt=1.0, dx=scene.dx, device=0,
                                     precision=U.PRECISION_BF16_CYCLE)
        ms_bf16 = lap_k.time_apply(0, 6, reps=args.kernel_reps)
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"kernel_only": True, "jacobi_ms": ms, "bytes_per_launch": lap.algorithmic_bytes(0, 2),
                          "jacobi_bf16_ms": ms_bf16, "bf16_bytes_per_launch": lap_k.algorithmic_bytes(0, 6),
                          "copy_bytes_each_way": src.numel() * 4, "copies": 10}), flush=True)
        return 0
    def timed_solves(l, x):
        """warm-up then timed solves from a zero guess; (ms/solve, latencies ms, iterations)"""
        it = None
        for _ in range(args.warmup):
            x.zero_()
            rc, it = l.solve_pcg_device(x.data_ptr(), b.data_ptr())
            if rc != U.SUCCESS:
                raise RuntimeError(f"warmup solve returned {rc}")
        torch.cuda.synchronize()
        lat = []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            ts = time.perf_counter()
            x.zero_()
            rc, it = l.solve_pcg_device(x.data_ptr(), b.data_ptr())
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - ts)
            if rc != U.SUCCESS:
                raise RuntimeError(f"solve returned {rc}")
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t_start) / args.steps, np.array(lat) * 1e3, it

    if args.profile_form:
        # a profiled run of one form's solves alone (run it with the level-to-kernel
        # mapping pinned, UAAMG_BLOCK_LEAVES, so that no timing launches enter the trace)
        lp = lap if args.profile_form == "fp32" else U.LaplacianWithLevel(
            scene.phi, scene.weights, dt=1.0, dx=scene.dx, device=0, precision=U.PRECISION_BF16_CYCLE)
        make_solver(lp)
        ms, lat, it = timed_solves(lp, x)
        print(json.dumps({"profile_form": args.profile_form, "n": args.n, "ms_per_solve": round(ms, 3),
                          "solves": args.warmup + args.steps, "pcg_iterations": it,
                          "kernel_map": lp.kernel_map()}), flush=True)
        return 0
    # the plain FP32 form (the reference's arithmetic)
    ms32, lat32, it32 = timed_solves(lap, x)
    x32 = x.cpu().numpy()
    # FP32 with the same FP64 residual replacement as the mixed form: the
    # like-for-like comparison at equal true-residual accuracy
    lap_rr = U.LaplacianWithLevel(scene.phi, scene.weights, dt=1.0, dx=scene.dx, device=0, flags=64)
    make_solver(lap_rr)
    x_rr = torch.zeros_like(b)
    ms_rr, lat_rr, it_rr = timed_solves(lap_rr, x_rr)
    true_nu_rr, reps_rr = lap_rr.last_residual()
    x_rr_h = x_rr.cpu().numpy()
    del lap_rr
    # the mixed-precision form: BF16-stored W-cycle vectors, FP64 residual replacement
    t0 = time.perf_counter()
    lap_mx = U.LaplacianWithLevel(scene.phi, scene.weights, dt=1.0, dx=scene.dx, device=0,
                                  precision=U.PRECISION_BF16_CYCLE)
    t_build_mx = time.perf_counter() - t0
    make_solver(lap_mx)
    x_mx = torch.zeros_like(b)
    ms_mx, lat_mx, it_mx = timed_solves(lap_mx, x_mx)
    true_nu_mx, reps_mx = lap_mx.last_residual()
    x_gpu = x_mx.cpu().numpy()
    ms_per_solve, lat_ms, iters = ms_mx, lat_mx, it_mx

    # dominant operator kernel of the measured (mixed) form: the level-0
    # weighted Jacobi sweep with BF16 x / y and the FP32 residual as b; the
    # FP32 form's sweep beside it
    def roof(l, mode, name):
        nbytes = l.algorithmic_bytes(0, mode)
        ms = l.time_apply(0, mode, reps=args.kernel_reps)
        ach = nbytes / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(name),
                "kernel": name, "bytes_per_launch": nbytes, "ms_per_launch": round(ms, 5)}

    if args.no_kernel_timing:   # a profiled run: only the solves' own launches in the trace
        roofline = roofline_fp32 = roof_sweep = None
        kernels_mx = kernels_32 = []
    else:
        roof_sweep = roof(lap_mx, 6, "apply_kernel<2, true, bf16, float, bf16> (level-0 Jacobi sweep, BF16 W-cycle)")
        roofline_fp32 = roof(lap, 2, "apply_kernel<2, true> (level-0 weighted Jacobi sweep, FP32)")
        # the kernel that takes the most time per solve (live launch times x launches per solve)
        kernels_mx = kernel_rooflines(lap_mx, it_mx, args.kernel_reps)
        kernels_32 = kernel_rooflines(lap, it32, args.kernel_reps)
        top = max(kernels_mx, key=lambda k: k["ms_per_solve"])
        roofline = roofline_entry(top, traffic_from_profiles(top["kernel"]))
    # the whole solve: algorithmic bytes of every operation / ms per solve
    bytes_mx = solve_bytes(lap_mx, it_mx, reps_mx, True)
    bytes_32 = solve_bytes(lap, it32, 0, False)
    roofline_solve = {"bound": "hbm", "achieved": round(bytes_mx / (ms_mx * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(bytes_mx / (ms_mx * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "bytes_per_solve": round(bytes_mx), "ms_per_solve": round(ms_mx, 4),
                      "fp32_plain": {"bytes_per_solve": round(bytes_32), "ms_per_solve": round(ms32, 4),
                                     "frac": round(bytes_32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}

    info = [lap.level_info(l) for l in range(lap.nlevels)]
    # one Jacobi sweep per level, back-to-back launches (the coarse levels show launch latency)
    level_sweep_us = ([] if args.no_kernel_timing else
                      [round(1e3 * lap.time_apply(l, 2, reps=100), 2) for l in range(lap.nlevels)])
    out = {
        "metric": METRIC,
        "value": round(ms_per_solve, 3),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_solve, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": "bf16 W-cycle level vectors; f32 PCG vectors, f64 dot products and residual replacement",
        "data": "synthetic",
        "config": {
            "workload": f"poisson {args.n}^3 config {args.config} (30% pool + 8 seeded spheres + solid sphere "
                        f"r=0.2), tol {args.tol:g}, mixed precision (BF16-stored W-cycle), smoother "
                        f"{args.smoother}",
            "n": args.n, "ndof": lap.mNumDof, "levels": lap.nlevels,
            "level_ndof": [i["ndof"] for i in info], "level_leaves": [i["nleaf_dof"] for i in info],
        },
        "solves_per_s": round(1e3 / ms_per_solve, 2),
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "pcg_iterations": iters,
        "residual_replacements": reps_mx,
        "fp32_plain": {"ms_per_solve": round(ms32, 3), "p50_ms": round(float(np.percentile(lat32, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat32, 99)), 3), "pcg_iterations": it32},
        "speedup_vs_fp32_plain": round(ms32 / ms_mx, 3),
        "fp32_replaced": {"ms_per_solve": round(ms_rr, 3), "p50_ms": round(float(np.percentile(lat_rr, 50)), 3),
                          "p99_ms": round(float(np.percentile(lat_rr, 99)), 3), "pcg_iterations": it_rr,
                          "residual_replacements": reps_rr, "true_residual_inf_fp64": true_nu_rr},
        "speedup_vs_fp32_replaced": round(ms_rr / ms_mx, 3),
        "level_sweep_us": level_sweep_us,
        "level_kernel_map": {"fp32": lap.kernel_map(), "bf16": lap_mx.kernel_map()},
        "build_s": round(t_build, 3),
        "build_s_mixed": round(t_build_mx, 3),
        "scene_s": round(t_scene, 3),
        "roofline": roofline,
        "roofline_solve": roofline_solve,
        "roofline_level0_sweep": roof_sweep,
        "roofline_fp32": roofline_fp32,
        "kernel_time_per_solve": {
            "bf16": [{"kernel": k["kernel"], "ms_per_launch": round(k["ms_per_launch"], 5),
                      "launches": k["launches_per_solve"], "ms_per_solve": round(k["ms_per_solve"], 4),
                      "frac": round(k["bytes_per_launch"] / (k["ms_per_launch"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                     for k in kernels_mx],
            "fp32": [{"kernel": k["kernel"], "ms_per_launch": round(k["ms_per_launch"], 5),
                      "launches": k["launches_per_solve"], "ms_per_solve": round(k["ms_per_solve"], 4),
                      "frac": round(k["bytes_per_launch"] / (k["ms_per_launch"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                     for k in kernels_32]},
    }

    out["rhs_build"] = rhs_build(lap, scene, b_host, t_rhs_host, dev)
    out["pressure_gradient"] = pressure_gradient(lap, scene, x_mx)
    out["face_weights"] = face_weights(scene)
    # handle creation (hierarchy build, upload, kernel-mapping tuning) with the level-0 per-voxel
    # stage on the device (the default) and on the host threads (flags bit 8), best of 3 each
    def build_time(flags):
        best = 1e9
        for _ in range(3):
            t0 = time.perf_counter()
            h = U.LaplacianWithLevel(scene.phi, scene.weights, dt=1.0, dx=scene.dx, device=0, flags=flags)
            best = min(best, time.perf_counter() - t0)
            del h
        return round(best, 3)
    out["handle_create_s"] = {"device_stage": build_time(0), "host_stage": build_time(256)}

    if not args.no_cpu_baseline:
        import oracle as O
        orc = O.OraclePoisson(scene.phi, scene.weights, 1.0, scene.dx)
        b_orc = orc.build_rhs(scene.weights, scene.vx, scene.vy, scene.vz)
        orc.solve_pcg(b_orc, tol=args.tol, smoother=args.smoother)   # warm-up
        cpu_times = []
        for _ in range(5):
            t0 = time.perf_counter()
            st_c, it_c, x_cpu = orc.solve_pcg(b_orc, tol=args.tol, smoother=args.smoother)
            cpu_times.append(time.perf_counter() - t0)
        cpu_s = float(np.median(cpu_times))
        st64, it64, x64 = orc.solve_pcg64(b_orc.astype(np.float64), tol=1e-12, smoother=args.smoother)
        r64 = b_orc.astype(np.float64) - orc.apply64(x_gpu.astype(np.float64))
        r64_32 = b_orc.astype(np.float64) - orc.apply64(x32.astype(np.float64))
        out["cpu_baseline"] = {"value": round(cpu_s * 1e3, 2), "unit": "ms/solve", "cores": O.get_threads(),
                               "kind": "port",
                               "sample": f"median of 5 FP32 solves (after one warm-up) of the same {args.n}^3 system "
                                         f"by the oracle, a C++ restatement of the reference's algorithm (not the "
                                         f"reference binary, which needs OpenVDB / Eigen / TBB), hierarchy build not "
                                         f"timed, {it_c} PCG iterations, OpenMP over leaves"}
        out["cpu_iterations"] = it_c
        out["rel_l2_err_vs_fp64_oracle"] = float(np.linalg.norm(x_gpu - x64) / np.linalg.norm(x64))
        out["rel_l2_vs_cpu_fp32"] = float(np.linalg.norm(x_gpu - x_cpu) / np.linalg.norm(x_cpu))
        out["true_rel_residual_inf"] = float(np.abs(r64).max() / np.abs(b_orc).max())
        out["fp32_plain"]["rel_l2_err_vs_fp64_oracle"] = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))
        out["fp32_plain"]["true_rel_residual_inf"] = float(np.abs(r64_32).max() / np.abs(b_orc).max())
        out["fp32_plain"]["rel_l2_vs_cpu_fp32"] = float(np.linalg.norm(x32 - x_cpu) / np.linalg.norm(x_cpu))
        r64_rr = b_orc.astype(np.float64) - orc.apply64(x_rr_h.astype(np.float64))
        out["fp32_replaced"]["rel_l2_err_vs_fp64_oracle"] = float(np.linalg.norm(x_rr_h - x64) / np.linalg.norm(x64))
        out["fp32_replaced"]["true_rel_residual_inf"] = float(np.abs(r64_rr).max() / np.abs(b_orc).max())
        out["speedup_vs_cpu_oracle_port"] = round(cpu_s * 1e3 / ms_per_solve, 2)
    if not args.no_512 and args.n != 512:
        out["poisson_512"] = run_poisson_512(args, dev)
    if not args.no_viscosity:
        out["viscosity"] = run_viscosity(args, dev)
    if not args.no_batch:
        out["batch"] = run_batch(args, dev)
        if not args.no_viscosity:
            out["viscosity_batch"] = run_visc_batch(args, dev)
    if not args.no_serve:
        out["serving"] = run_serving(args, dev)
        out["serving_tight_budget"] = run_tight_serving(args, dev, out["serving"])
    print(json.dumps(out), flush=True)
    return 0


def rhs_build(lap, scene, b_host, t_host_s: float, dev) -> dict:
    """The right-hand side built on the host (uaamg_poisson_build_rhs) and on the
    device (uaamg_poisson_build_rhs_device) from host-resident grids (values
    uploaded in the call) and from device-resident grid values; the device
    results are checked bit for bit against the host's."""
    import numpy as np
    import torch

    grids = [scene.weights, scene.vx, scene.vy, scene.vz]
    out = torch.empty(lap.mNumDof, dtype=torch.float32, device=dev)

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / reps

    ms_upload = timed(lambda: lap.createPressurePoissonRightHandSide_device(out.data_ptr(), *grids))
    same_upload = bool(np.array_equal(out.cpu().numpy(), b_host))
    dvals = [torch.from_numpy(np.ascontiguousarray(g.values, dtype=np.float32)).to(dev) for g in grids]
    ptrs = [t.data_ptr() for t in dvals]
    out.zero_()
    ms_resident = timed(lambda: lap.createPressurePoissonRightHandSide_device(out.data_ptr(), *grids, values_ptrs=ptrs))
    same_resident = bool(np.array_equal(out.cpu().numpy(), b_host))
    return {"host_ms": round(1e3 * t_host_s, 2), "device_ms_values_uploaded": round(ms_upload, 2),
            "device_ms_values_resident": round(ms_resident, 2),
            "input_mb": round(sum(g.values.nbytes for g in grids) / 1e6, 1),
            "identical_to_host": same_upload and same_resident}


def pressure_gradient(lap, scene, pressure) -> dict:
    """The step after the solve on the device (uaamg_poisson_apply_pressure_gradient_device) on the
    solved pressure, velocity grids in host memory (uploaded and downloaded in the call): ms per call
    and the number of faces it updated."""
    import numpy as np
    import torch

    vel = [scene.vx, scene.vy, scene.vz]
    lap.apply_pressure_gradient_device(pressure.data_ptr(), scene.phi, scene.weights, None, vel, 0.04, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = lap.apply_pressure_gradient_device(pressure.data_ptr(), scene.phi, scene.weights, None, vel, 0.04, 0.0)
    ms = 1e3 * (time.perf_counter() - t0)
    faces = int(sum(np.unpackbits(a.view(np.uint8)).sum() for a, _, _, _ in out))
    mb = sum(g.values.nbytes for g in vel) / 1e6
    # the device-resident calling form: every values array already in HBM, the call reads and writes
    # them there (the leaf tables and masks go through the handle's reusable buffer: no allocation)
    import ctypes as C

    import uaamg_amd as U
    dev = pressure.device
    keep = []

    def on_dev(a):
        t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
        keep.append(t)
        return t.data_ptr()
    st = [g._struct(on_dev(g.values)) for g in (scene.phi, scene.weights)]
    chans = (U._FaceChannel * 3)()
    masks, orgs = [], []
    for c, g in enumerate(vel):
        org = np.ascontiguousarray(g.origins, dtype=np.int32)
        n = org.shape[0]
        act0 = (np.full((n, 8), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64) if g.active is None
                else np.array(g.active, dtype=np.uint64, copy=True))
        act = act0.copy()
        masks.append((act0, act))
        orgs.append(org)
        chans[c].nleaf = n
        chans[c].origins = org.ctypes.data
        chans[c].active = act.ctypes.data
        chans[c].values = on_dev(g.values)
        chans[c].update = on_dev(np.zeros_like(g.values))
        chans[c].reflected = on_dev(np.zeros_like(g.values))
    torch.cuda.synchronize()

    def call():
        for act0, act in masks:
            act[...] = act0
        rc = U.lib().uaamg_poisson_apply_pressure_gradient_device(lap._h, C.c_void_p(pressure.data_ptr()),
                                                                   C.byref(st[0]), C.byref(st[1]), None, chans,
                                                                   0.04, 0.0, 1)
        if rc != 0:
            raise RuntimeError(f"pressure gradient returned {rc}")
    call()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    torch.cuda.synchronize()
    ms_dev = 1e3 * (time.perf_counter() - t0) / reps
    del keep
    return {"ms_per_call": round(ms, 2), "faces_updated": faces, "velocity_mb": round(mb, 1),
            "io": "velocity, update and reflected grids through host memory",
            "ms_per_call_values_resident": round(ms_dev, 2),
            "io_values_resident": "velocity, update, reflected, phi and weights already in HBM; origins and the "
                                  "active masks (in/out) through host memory; no device allocation per call"}


def face_weights(scene) -> dict:
    """Face weights from a solid SDF on the device (uaamg_face_weights): the scene's liquid SDF
    topology dilated twice, a solid sphere SDF sampled on the weight grid's leaves; ms per call
    including the host topology pass and the transfers."""
    import numpy as np

    import uaamg_amd as U

    org = scene.weights.origins
    n = int(round(1.0 / scene.dx))
    off = np.arange(512)
    lx, ly, lz = off >> 6, (off >> 3) & 7, off & 7
    vals = np.empty((org.shape[0], 512), np.float32)
    for i, o in enumerate(org):
        vals[i] = np.sqrt((o[0] + lx - 0.5 * n) ** 2 + (o[1] + ly - 0.5 * n) ** 2 + (o[2] + lz - 0.5 * n) ** 2) - 0.2 * n
    solid = U.Grid(org, vals, None, (3.0,))
    U.calculate_face_weights(scene.phi, solid)
    t0 = time.perf_counter()
    g = U.calculate_face_weights(scene.phi, solid)
    ms = 1e3 * (time.perf_counter() - t0)
    return {"ms_per_call": round(ms, 2), "leaves": int(g.origins.shape[0]),
            "io": "liquid topology and solid SDF from host memory, weights back to host memory"}


def run_serving(args, dev, budget_bytes: int | None = None, seconds: float | None = None) -> dict:
    """The serving configuration through the C++ solve server: a pool of
    problems (80% Poisson with edge log-uniform in [48, 384], 20% viscosity
    with edge log-uniform in [48, 192]) registered once, requests arriving as a
    Poisson process, each re-solving one problem warm-started from its resident
    previous solution with a right-hand side perturbed by 5% (pre-generated on
    the device). The server admits requests as its `--serve-workers` slots and
    its memory budget allow; problems own their streams, so admitted solves
    overlap on the device. Latency is the server's own submission-to-completion
    time; p50 / p99 and completed solves/s over the run. budget_bytes: the
    server's memory budget (default: --serve-budget-gb, 0 = 90% of free)."""
    import threading

    import numpy as np
    import torch

    import uaamg_amd as U
    from uaamg_amd.server import SolveServer

    rng = np.random.default_rng(2024)
    n_pool, n_visc, n_rhs = 10, 2, 4
    budget = int(args.serve_budget_gb * (1 << 30)) if args.serve_budget_gb > 0 else 0
    if budget_bytes is not None:
        budget = int(budget_bytes)
    run_seconds = args.serve_seconds if seconds is None else seconds
    srv = SolveServer(device=0, slots=args.serve_workers, memory_budget=budget)
    t0 = time.perf_counter()
    pool = []
    for k in range(n_pool):
        visc = k < n_visc
        lo, hi = (48, 192) if visc else (48, 384)
        n = int(round(lo * np.exp(rng.uniform() * np.log(hi / lo))))
        n = max(lo, min(hi, 8 * round(n / 8)))
        if visc:
            sc = U.ViscScene(3, n, seed=100 + k, mu=1.0)
            lw = U.L_with_level(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx, device=-1)
            b0 = torch.from_numpy(lw.build_rhs(sc.velocity).data).to(dev)
            pid = srv.add_viscosity(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx)
            tol = 1e-6
        else:
            sc = U.Scene(3, n, seed=100 + k)
            lap = U.LaplacianWithLevel(sc.phi, sc.weights, dt=1.0, dx=sc.dx, device=-1)
            b0 = torch.from_numpy(lap.createPressurePoissonRightHandSide(sc.weights, sc.vx, sc.vy, sc.vz)).to(dev)
            pid = srv.add_poisson(sc.phi, sc.weights, 1.0, sc.dx, precision=args.serve_precision)
            tol = args.tol
        g = torch.Generator(device=dev).manual_seed(k)
        rhs = [b0 * (1 + 0.05 * torch.randn(b0.shape, device=dev, generator=g)) for _ in range(n_rhs)]
        pool.append({"n": n, "visc": visc, "pid": pid, "rhs": rhs, "tol": tol, "ndof": int(b0.numel())})
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    # warm-up: one solve each (residency, graph capture) and the mean service time
    svc = []
    for p in pool:
        st, it, qms, tms = srv.wait(srv.submit(p["pid"], p["rhs"][0].data_ptr(), tol=p["tol"]))
        if st < 0:
            raise RuntimeError(f"serving warm-up solve returned {st}")
        svc.append(tms * 1e-3)
    mean_svc = float(np.mean(svc))
    rate = 1.0 / mean_svc   # one request per mean service time
    srv_bytes = [int(srv.problem_info(p["pid"])["device_bytes"]) for p in pool]
    base = srv.stats()

    tickets, lock, done = [], threading.Lock(), threading.Event()
    results, req_pool, errors = [], [], []

    def collector():
        k = 0
        while True:
            with lock:
                t = tickets[k] if k < len(tickets) else None
            if t is None:
                if done.is_set():
                    with lock:
                        if k >= len(tickets):
                            return
                    continue
                time.sleep(1e-4)
                continue
            r = srv.wait(t)
            results.append(r)
            if r[0] < 0:
                errors.append(srv.last_error())
            k += 1

    col = threading.Thread(target=collector, daemon=True)
    col.start()
    t_start = time.perf_counter()
    t_next = t_start
    sent = 0
    while True:
        t_next += rng.exponential(1.0 / rate)
        if t_next - t_start > run_seconds:
            break
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        k_pool = int(rng.integers(n_pool))
        p = pool[k_pool]
        t = srv.submit(p["pid"], p["rhs"][int(rng.integers(n_rhs))].data_ptr(), tol=p["tol"])
        with lock:
            tickets.append(t)
            req_pool.append(k_pool)
        sent += 1
    done.set()
    col.join()
    wall = time.perf_counter() - t_start
    st = srv.stats()
    srv.close()
    lat_ms = np.array([r[3] for r in results])
    fails = sum(1 for r in results if r[0] != 0)
    # non-SUCCESS results by status and problem (status 1 = FAILED, the reference's
    # PCG failure rule; negative = a server or device error)
    fail_detail = {}
    for r, k in zip(results, req_pool):
        if r[0] != 0:
            key = f"status {r[0]} / {'viscosity' if pool[k]['visc'] else 'poisson'} n={pool[k]['n']}"
            fail_detail[key] = fail_detail.get(key, 0) + 1
    return {"metric": "p50 / p99 latency (submission to completion) and solves/s, Poisson-process arrivals",
            "engine": "C++ solve server (uaamg_server_*), continuous admission",
            "requests": sent, "completed": len(results), "failed": fails, "failed_by_status": fail_detail,
            "errors": sorted(set(errors))[:8],
            "solves_per_s": round(len(results) / wall, 2),
            "arrival_rate_per_s": round(rate, 2), "slots": args.serve_workers,
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 2) if len(results) else None,
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 2) if len(results) else None,
            "mean_queue_ms": round(float(np.mean([r[2] for r in results])), 3) if len(results) else None,
            "mean_service_ms": round(1e3 * mean_svc, 2),
            "peak_inflight": st["peak_inflight"], "evictions": st["evictions"] - base["evictions"],
            "resident_bytes": st["resident_bytes"], "memory_budget_bytes": budget or "90% of free",
            "largest_problem_bytes": max(srv_bytes) if srv_bytes else 0,
            "poisson_precision": "bf16 W-cycle" if args.serve_precision == 1 else "fp32",
            "pool": [{"n": p["n"], "kind": "viscosity" if p["visc"] else "poisson", "ndof": p["ndof"]}
                     for p in pool],
            "build_s": round(t_build, 2)}


def run_tight_serving(args, dev, first: dict) -> dict:
    """The same stream under a memory budget of half the pool's resident footprint: admission waits
    for memory and idle problems are evicted (least recently used first) and restored."""
    tight = max(int(0.5 * first["resident_bytes"]), int(1.05 * first["largest_problem_bytes"]))
    return run_serving(args, dev, budget_bytes=tight, seconds=0.5 * args.serve_seconds)


def run_poisson_512(args, dev) -> dict:
    """Config 1 at 512^3 (about 47M DOF): ms per solve of both forms, b in HBM."""
    import numpy as np
    import torch

    import uaamg_amd as U

    sc = U.Scene(args.config, 512, seed=args.seed)
    out = {"n": 512}
    b = None
    for prec, key in ((U.PRECISION_FP32, "fp32_plain"), (U.PRECISION_BF16_CYCLE, "mixed")):
        lap = U.LaplacianWithLevel(sc.phi, sc.weights, dt=1.0, dx=sc.dx, device=0, precision=prec)
        s = U.PoissonSolver(lap)
        s.mRelativeTolerance = args.tol
        s._push_options()
        if b is None:
            b = torch.from_numpy(lap.createPressurePoissonRightHandSide(sc.weights, sc.vx, sc.vy, sc.vz)).to(dev)
            out["ndof"] = lap.mNumDof
            out["levels"] = lap.nlevels
        x = torch.zeros_like(b)
        lap.solve_pcg_device(x.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        lat = []
        for _ in range(3):
            x.zero_()
            t = time.perf_counter()
            rc, it = lap.solve_pcg_device(x.data_ptr(), b.data_ptr())
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t)
            if rc != U.SUCCESS:
                raise RuntimeError(f"512^3 solve returned {rc}")
        out[key] = {"ms_per_solve": round(1e3 * float(np.mean(lat)), 3), "pcg_iterations": it,
                    "true_rel_residual_inf_fp64": (lap.last_residual()[0] / float(b.abs().max().item())
                                                   if prec == U.PRECISION_BF16_CYCLE else None)}
        del lap
    out["speedup_vs_fp32_plain"] = round(out["fp32_plain"]["ms_per_solve"] / out["mixed"]["ms_per_solve"], 3)
    return out


def run_batch(args, dev) -> dict:
    """The batch configuration: 64 independent Poisson problems (edge
    log-uniform in [32, 192], 1-6 spheres over a 10-50% pool, solid sphere
    with probability 1/2; seeds 0..63), tol 1e-7, b resident in HBM. Reports
    solves/s for: the problems one after another (one handle each), the
    stream-per-problem entry point (`--batch-concurrency` host threads), and
    the merged batch (one leaf table per level, one launch per W-cycle stage
    for all problems) in FP32 and with the BF16 W-cycle. The merged FP32
    results must equal the sequential ones bit for bit."""
    import numpy as np
    import torch

    import uaamg_amd as U

    t0 = time.perf_counter()
    scenes, laps, xs, bs, ndof = [], [], [], [], 0
    for seed in range(args.batch_size):
        sc = U.Scene(3, 0, seed=seed)
        lap = U.LaplacianWithLevel(sc.phi, sc.weights, dt=1.0, dx=sc.dx, device=0)
        s = U.PoissonSolver(lap)
        s.mRelativeTolerance = 1e-7
        s._push_options()
        b = torch.from_numpy(lap.createPressurePoissonRightHandSide(sc.weights, sc.vx, sc.vy, sc.vz)).to(dev)
        scenes.append(sc)
        laps.append(lap)
        bs.append(b)
        xs.append(torch.zeros_like(b))
        ndof += lap.mNumDof
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    merged = U.PoissonBatch([(sc.phi, sc.weights, 1.0, sc.dx) for sc in scenes], precision=0)
    t_build_merged = time.perf_counter() - t0
    merged_mx = U.PoissonBatch([(sc.phi, sc.weights, 1.0, sc.dx) for sc in scenes], precision=1)

    def sequential():
        its = []
        for lap, x, b in zip(laps, xs, bs):
            x.zero_()
            rc, it = lap.solve_pcg_device(x.data_ptr(), b.data_ptr())
            its.append((rc, it))
        torch.cuda.synchronize()
        return its

    def streams():
        for x in xs:
            x.zero_()
        torch.cuda.synchronize()
        st, its = U.solve_batch_device(laps, [x.data_ptr() for x in xs], [b.data_ptr() for b in bs],
                                       args.batch_concurrency)
        torch.cuda.synchronize()
        return list(zip(st, its))

    def merged_solve(bt):
        def f():
            for x in xs:
                x.zero_()
            torch.cuda.synchronize()
            st, its = bt.solve_device([x.data_ptr() for x in xs], [b.data_ptr() for b in bs])
            torch.cuda.synchronize()
            return list(zip(st, its))
        return f

    def timed(fn, reps=3):
        fn()
        t = time.perf_counter()
        for _ in range(reps):
            its = fn()
        return (time.perf_counter() - t) / reps, its

    t_seq, its_seq = timed(sequential)
    x_seq = [x.cpu().numpy() for x in xs]
    t_str, its_str = timed(streams)
    same_str = all(np.array_equal(a, x.cpu().numpy()) for a, x in zip(x_seq, xs)) and its_seq == its_str
    t_mrg, its_mrg = timed(merged_solve(merged))
    same_mrg = all(np.array_equal(a, x.cpu().numpy()) for a, x in zip(x_seq, xs)) and its_seq == its_mrg
    t_mx, its_mx = timed(merged_solve(merged_mx))
    x_mx = [x.cpu().numpy() for x in xs]
    dev_mx = max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)) for a, b in zip(x_mx, x_seq))
    n = args.batch_size
    info = merged.info()
    # solves/s count the solves that met the tolerance (SUCCESS); a FAILED one (the reference's rule
    # stops it) is reported, not counted
    ok_mx = sum(1 for st, _ in its_mx if st == U.SUCCESS)
    ok_32 = sum(1 for st, _ in its_mrg if st == U.SUCCESS)
    ok_seq = sum(1 for st, _ in its_seq if st == U.SUCCESS)
    return {"metric": f"solves/s over {n} independent Poisson problems (edge log-uniform [32,192]), tol 1e-7, "
                      f"SUCCESS solves only",
            "value": round(ok_mx / t_mx, 1), "unit": "solves/s", "form": "merged batch, BF16 W-cycle",
            "problems": n, "total_ndof": ndof,
            "ms_per_batch": round(1e3 * t_mx, 2), "succeeded_bf16": ok_mx, "succeeded_fp32": ok_32,
            "merged_fp32_ms_per_batch": round(1e3 * t_mrg, 2), "merged_fp32_solves_per_s": round(ok_32 / t_mrg, 1),
            "streams_ms_per_batch": round(1e3 * t_str, 2), "streams_solves_per_s": round(ok_seq / t_str, 1),
            "sequential_ms_per_batch": round(1e3 * t_seq, 2), "sequential_solves_per_s": round(ok_seq / t_seq, 1),
            "concurrency": args.batch_concurrency, "merged_levels": info["nlevels"],
            "merged_leaves0": info["leaves0"], "merged_device_bytes": info["device_bytes"],
            "mean_iterations": round(float(np.mean([i for _, i in its_mrg])), 2),
            "mean_iterations_bf16": round(float(np.mean([i for _, i in its_mx])), 2),
            "failed_fp32": [k for k, (st, _) in enumerate(its_mrg) if st != U.SUCCESS],
            "failed_bf16": [k for k, (st, _) in enumerate(its_mx) if st != U.SUCCESS],
            "merged_fp32_identical_to_sequential": bool(same_mrg),
            "streams_identical_to_sequential": bool(same_str),
            "bf16_max_rel_l2_vs_fp32": dev_mx,
            "build_s": round(t_build, 2), "merge_build_s": round(t_build_merged, 2)}


def run_visc_batch(args, dev) -> dict:
    """The viscosity batch: `--visc-batch-size` independent viscosity problems
    (the batch scenes: edge log-uniform in [32, 192], 1-6 spheres over a
    10-50% pool, solid sphere with probability 1/2; seeds 0..n-1; mu = 1,
    dt = 1/24, rho = 1000), tol 1e-6, b resident in HBM. Reports solves/s for
    the problems one after another (one handle each) and for the merged batch
    (one leaf table per level, one launch per colour pass for all problems),
    whose results must equal the sequential ones bit for bit."""
    import numpy as np
    import torch

    import uaamg_amd as U
    from uaamg_amd.viscosity import ViscBatch

    tol = 1e-6
    t0 = time.perf_counter()
    scenes, solvers, xs, bs, ndof = [], [], [], [], 0
    for seed in range(args.visc_batch_size):
        sc = U.ViscScene(3, 0, seed=seed, mu=1.0, dt=1.0 / 24.0, rho=1000.0)
        lw = U.L_with_level(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx, device=0)
        sv = U.simd_viscosity3d(lw, sc.velocity)
        b = torch.from_numpy(sv.m_rhs.data).to(dev)
        scenes.append(sc)
        solvers.append(sv)
        bs.append(b)
        xs.append(torch.zeros_like(b))
        ndof += lw.m_ndof
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    merged = ViscBatch([(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx) for sc in scenes])
    t_build_merged = time.perf_counter() - t0

    def sequential():
        its = []
        for sv, x, b in zip(solvers, xs, bs):
            x.zero_()
            its.append(sv.solve_pcg_device(tol, x.data_ptr(), b.data_ptr()))
        torch.cuda.synchronize()
        return its

    def merged_solve():
        for x in xs:
            x.zero_()
        torch.cuda.synchronize()
        its = merged.solve_device(tol, [x.data_ptr() for x in xs], [b.data_ptr() for b in bs])
        torch.cuda.synchronize()
        return its

    def timed(fn, reps=3):
        fn()
        t = time.perf_counter()
        for _ in range(reps):
            its = fn()
        return (time.perf_counter() - t) / reps, its

    t_seq, its_seq = timed(sequential)
    x_seq = [x.cpu().numpy() for x in xs]
    t_mrg, its_mrg = timed(merged_solve)
    same = all(np.array_equal(a, x.cpu().numpy()) for a, x in zip(x_seq, xs)) and its_seq == its_mrg
    n = args.visc_batch_size
    info = merged.info()
    return {"metric": f"solves/s over {n} independent viscosity problems (edge log-uniform [32,192]), tol 1e-6",
            "value": round(n / t_mrg, 1), "unit": "solves/s", "form": "merged batch, FP32",
            "problems": n, "total_ndof": ndof, "ms_per_batch": round(1e3 * t_mrg, 2),
            "sequential_ms_per_batch": round(1e3 * t_seq, 2), "sequential_solves_per_s": round(n / t_seq, 1),
            "speedup_vs_sequential": round(t_seq / t_mrg, 2),
            "merged_levels": info["nlevels"], "merged_leaves0": info["leaves0"],
            "merged_device_bytes": info["device_bytes"],
            "mean_iterations": round(float(np.mean(its_mrg)), 2),
            "merged_identical_to_sequential": bool(same),
            "build_s": round(t_build, 2), "merge_build_s": round(t_build_merged, 2)}


def run_viscosity(args, dev) -> dict:
    """The variational viscosity solve on the 128^3 pool + sphere scene (mu = 1,
    dt = 1/24, rho = 1000, tol 1e-6, zero initial guess, zero solid velocity):
    ms per pcg_solve with b resident in HBM, the dominant kernel's roofline
    (one red Gauss-Seidel pass over channel 0 of level 0) and the CPU oracle
    solving the same system."""
    import numpy as np
    import torch

    import uaamg_amd as U

    n, tol = args.visc_n, 1e-6
    t0 = time.perf_counter()
    sc = U.ViscScene(0, n, seed=0, mu=1.0, dt=1.0 / 24.0, rho=1000.0)
    t_scene = time.perf_counter() - t0
    t0 = time.perf_counter()
    lw = U.L_with_level(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx, device=0)
    t_build = time.perf_counter() - t0
    solver = U.simd_viscosity3d(lw, sc.velocity)
    if args.kernel_only:
        ms = lw.time_apply(0, 4, 0, reps=args.kernel_reps)
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        return {"kernel_only": True, "red_pass_ms": ms, "bytes_per_launch": lw.algorithmic_bytes(0, 4, 0),
                "copy_bytes_each_way": src.numel() * 4, "copies": 10}
    b = torch.from_numpy(solver.m_rhs.data).to(dev)
    x = torch.zeros_like(b)
    for _ in range(2):
        x.zero_()
        solver.solve_pcg_device(tol, x.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.visc_steps):
        ts = time.perf_counter()
        x.zero_()
        it = solver.solve_pcg_device(tol, x.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - ts)
    ms = 1e3 * (time.perf_counter() - t_start) / args.visc_steps
    x_gpu = x.cpu().numpy()
    # the mixed form: the W-cycle's level vectors below level 0 stored in BF16
    lw_mx = U.L_with_level(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx, device=0, precision=1)
    solver_mx = U.simd_viscosity3d(lw_mx, sc.velocity)
    for _ in range(2):
        x.zero_()
        solver_mx.solve_pcg_device(tol, x.data_ptr(), b.data_ptr())
    torch.cuda.synchronize()
    lat_mx = []
    t_start = time.perf_counter()
    for _ in range(args.visc_steps):
        ts = time.perf_counter()
        x.zero_()
        it_mx = solver_mx.solve_pcg_device(tol, x.data_ptr(), b.data_ptr())
        torch.cuda.synchronize()
        lat_mx.append(time.perf_counter() - ts)
    ms_mx = 1e3 * (time.perf_counter() - t_start) / args.visc_steps
    x_mx = x.cpu().numpy()
    map_mx = lw_mx.kernel_map()
    del solver_mx, lw_mx
    # handle creation with the level-0 volume stage on the device (the default) and on the host threads
    # (flags bit 8), best of 2 each
    def visc_build_time(flags):
        best = 1e9
        for _ in range(2):
            t0 = time.perf_counter()
            h = U.L_with_level(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx, device=0, flags=flags)
            best = min(best, time.perf_counter() - t0)
            del h
        return round(best, 3)
    create_s = {"device_volume_stage": visc_build_time(0), "host_volume_stage": visc_build_time(256)}
    red_bytes = lw.algorithmic_bytes(0, 4, 0)
    red_ms = lw.time_apply(0, 4, 0, reps=args.kernel_reps)
    ach = red_bytes / (red_ms * 1e-3) / 1e9
    info = [lw.level_info(l) for l in range(lw.nlevels)]
    out = {
        "metric": "ms/viscosity pcg_solve to 1e-6 @128^3 (pool 40% + sphere r=0.15, mu=1, dt=1/24, rho=1000)",
        "value": round(ms, 3), "unit": "ms/solve", "dtype": "f32",
        "ndof": lw.m_ndof, "levels": lw.nlevels, "level_ndof": [i["ndof"] for i in info],
        "iterations": it, "p50_ms": round(float(np.percentile(np.array(lat) * 1e3, 50)), 3),
        "p99_ms": round(float(np.percentile(np.array(lat) * 1e3, 99)), 3),
        "build_s": round(t_build, 3), "scene_s": round(t_scene, 3),
        "mixed": {"ms_per_solve": round(ms_mx, 3), "iterations": it_mx,
                  "p50_ms": round(float(np.percentile(np.array(lat_mx) * 1e3, 50)), 3),
                  "p99_ms": round(float(np.percentile(np.array(lat_mx) * 1e3, 99)), 3),
                  "storage": "bf16 W-cycle level vectors below level 0; f32 level 0 and PCG, f64 dot products",
                  "level_kernel_map": map_mx},
        "speedup_mixed_vs_fp32_plain": round(ms / ms_mx, 3),
        "handle_create_s": create_s,
        "level_kernel_map": lw.kernel_map(),
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles("visc_apply_kernel<4, 0> (level-0 red Gauss-Seidel pass, "
                                                      "channel 0)"),
                     "kernel": "visc_apply_kernel<4, 0> (level-0 red Gauss-Seidel pass, channel 0)",
                     "bytes_per_launch": red_bytes, "ms_per_launch": round(red_ms, 5)},
    }
    if not args.no_cpu_baseline:
        import oracle as O
        orc = O.OracleViscosity(sc.viscosity, sc.liquid, sc.solid, sc.dt, sc.rho, sc.dx)
        b_o = orc.build_rhs(sc.velocity, None)
        orc.solve(b_o, tol=tol)   # warm-up
        cpu_times = []
        for _ in range(5):
            t0 = time.perf_counter()
            it_c, x_cpu, _ = orc.solve(b_o, tol=tol)
            cpu_times.append(time.perf_counter() - t0)
        cpu_s = float(np.median(cpu_times))
        it64, x64 = orc.solve64(b_o.astype(np.float64), tol=1e-12)
        out["cpu_baseline"] = {"value": round(cpu_s * 1e3, 2), "unit": "ms/solve", "cores": O.get_threads(),
                               "kind": "port",
                               "sample": f"median of 5 FP32 pcg_solves (after one warm-up) of the same {n}^3 system by "
                                         f"the oracle, a C++ restatement of the reference's algorithm, hierarchy build "
                                         f"not timed, {it_c} iterations, OpenMP over leaves"}
        out["cpu_iterations"] = it_c
        out["rel_l2_err_vs_fp64_oracle"] = float(np.linalg.norm(x_gpu - x64) / np.linalg.norm(x64))
        out["rel_l2_vs_cpu_fp32"] = float(np.linalg.norm(x_gpu - x_cpu) / np.linalg.norm(x_cpu))
        out["mixed"]["rel_l2_err_vs_fp64_oracle"] = float(np.linalg.norm(x_mx - x64) / np.linalg.norm(x64))
        out["speedup_vs_cpu_oracle_port"] = round(cpu_s * 1e3 / ms, 2)
    return out


if __name__ == "__main__":
    sys.exit(main())
