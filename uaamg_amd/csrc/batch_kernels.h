// Kernels of the batched Poisson solve: many independent problems whose
// leaves are merged into one table per level, so each W-cycle stage is one
// launch for the whole batch.
//
// Level 0 of the merged table gives every problem a leaf range that starts at
// a multiple of kWavesPerBlock leaves and is padded to one (padding leaves have
// no DOF), so a reduction block never straddles two problems and problem p's
// blocks [blk_begin[p], blk_begin[p+1]) are exactly the blocks a solve of p
// alone would form, in the same order. The per-problem finalize then sums them
// in the single-problem finalize's order, which makes every scalar - and so
// every iterate - bit-identical to solving each problem on its own.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "poisson_kernels.h"

namespace uaamg {

// Per-problem outer-PCG state in device memory.
struct BatchPcg {
    double rho, sigma, rho_new;
    float alpha, beta;
    float nu, nu_prev, numax;
    int32_t state;         // BatchState
    int32_t iters;         // PCG iterations taken (mIterationTaken)
    int32_t max_iters;
    float tol;
    int32_t it0;           // iterations at the start of the current recurrence (growth rule)
    float true_nu;         // FP64 true residual of the last replacement check (-1 without)
    float prev_true;
    int32_t replacements;  // residuals replaced by the true residual: steps and restarts (-1 without replacement)
    int32_t refine;        // 1 while the replacement step may still restart this problem
    int32_t restarts;      // PCG restarts from the true residual
    int32_t replacing;     // the next step forms r as the true residual of its iterate
    int32_t stalled;       // the true residual stopped halving in this recurrence
};

enum BatchState : int32_t { kBActive = 0, kBConverging = 1, kBSuccess = 2, kBFailed = 3 };

// finalize operations of the batch
enum BatchFinOp : int {
    kBFinNu0 = 0,
    kBFinRho = 1,
    kBFinRhoNew = 2,
    kBFinSigma = 3,
    kBFinNuFirst = 4,
    kBFinNuIter = 5,
    kBFinReplace = 6,   // FP64 true residual: restart problems whose true residual misses the tolerance
    // the viscosity pcg_solve rule (simd_viscosity3d.cpp:3694-3740): stop when ||r||_inf <= tol ||r0||_inf
    // or at the last allowed iteration, no failure test; the iteration form counts the step first
    kBFinNuViscFirst = 7,
    kBFinNuViscIter = 8,
};
constexpr int kBMaxReplacements = 4;

// partials of the level-0 reductions, one per block (no finalize)
hipError_t launch_absmax_partials(const LevelView& L, const float* r, double* partials, hipStream_t s);
// r = b - A x in FP64 (stored FP32) and its partial max|r|
hipError_t launch_residual64_partials(const LevelView& L, const float* x, const float* b, float* r, double* partials,
                                      hipStream_t s);
hipError_t launch_dot_partials(const LevelView& L, const float* a, const float* b, double* partials, hipStream_t s);
hipError_t launch_laplacian_dot_partials(const LevelView& L, const float* p, float* q, double* partials,
                                         hipStream_t s);
hipError_t launch_jacobi_dot_partials(const LevelView& L, const float* x, const float* b, float* y, float w,
                                      double* partials, hipStream_t s);
hipError_t launch_jacobi_dot_partials(const LevelView& L, const bf16* x, const float* b, float* y, float w,
                                      double* partials, hipStream_t s);

// Per-problem forms of the PCG vector kernels on the merged level 0
// (leafprob[leaf] = problem; only problems in kBActive are updated, a problem
// in kBConverging gets its last x += alpha p).
hipError_t launch_bupdate_laplacian_dot(const LevelView& L, const int32_t* leafprob, const BatchPcg* st,
                                        const float* z, const float* p, float* p_new, float* x, float* q,
                                        double* partials, hipStream_t s);
// r -= alpha q, or for problems with `replacing` set r = b - A (x + alpha p)
// in FP64 (the single-problem replacement step); partial max|r|
hipError_t launch_baxpy_absmax(const LevelView& L, const int32_t* leafprob, const BatchPcg* st, const float* q,
                               const float* x, const float* p, const float* b, float* r, double* partials,
                               hipStream_t s);
hipError_t launch_bupdate_x(const LevelView& L, const int32_t* leafprob, BatchPcg* st, const float* p, float* x,
                            hipStream_t s);
// x += alpha p; p = z + beta p for the problems in kBActive (update_kernel<true>'s arithmetic)
hipError_t launch_bupdate_xp(const LevelView& L, const int32_t* leafprob, const BatchPcg* st, const float* z,
                             float* x, float* p, hipStream_t s);
// One workgroup per problem: sums its blocks' partials in the single-problem
// order and applies `op` (scalars, then the convergence / failure rules);
// op < 0 moves problems whose last x update ran from kBConverging to kBSuccess.
hipError_t launch_bfinalize(int op, int nprob, const int32_t* blk_begin, const double* partials, BatchPcg* st,
                            hipStream_t s);
// Publishes (sequence << 32 | number of problems still iterating) to the mailbox.
hipError_t launch_bstatus(int nprob, const BatchPcg* st, unsigned int* seq, HostMailbox* mb, hipStream_t s);
// Coarsest solves of a group of problems, one workgroup each.
hipError_t launch_bcoarsest(const CoarsestView* views, int n, int max_ndof, const float* b, float* x,
                            hipStream_t s);
hipError_t launch_bcoarsest(const CoarsestView* views, int n, int max_ndof, const bf16* b, bf16* x,
                            hipStream_t s);

}  // namespace uaamg
