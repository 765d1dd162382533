// Coarsest-level solve as a block-wide device routine: Eigen::ConjugateGradient
// with DiagonalPreconditioner, zero initial guess, maxIterations = 10,
// tolerance = float epsilon, restated (constructCoarsestLevelExactSolver,
// simd_vdb_poisson.cpp:2114-2126; Eigen ConjugateGradient.h conjugate_gradient()).
//
// Each thread keeps up to kPer DOFs' matrix rows and x, r, z, p, q = A p in
// registers; only z is shared through LDS for the neighbour reads. An absent
// neighbour reads slot kMaxCoarsestDof of the shared vector, which holds
// zero, so the six reads of a row are unconditional. Two barriers per
// iteration: the reduction of <p, Ap>, and the reduction of ||r||^2 and
// <r, z>, whose barrier also publishes the new z. The new direction's
// product is then formed as A p = A z + beta A p_old (the same vector as
// Eigen's A p, rounded differently), so no third barrier publishes p.
#pragma once

#include <cfloat>

#include "poisson_kernels.h"
#include "wave_reduce.h"

namespace uaamg {
namespace dev {

template <int kThreads>
struct CgShared {
    float sp[kMaxCoarsestDof + 1];   // the last slot stays zero
    float red[2][kThreads / 64][2];
};

__device__ __forceinline__ float cg_wave_sum(float v) { return wave_sum_dpp(v); }

// Sums (a, b) over the block with one barrier; red alternates between two
// buffers, so the next call needs no barrier before it writes.
template <int kThreads>
__device__ __forceinline__ float2 cg_sum2(float a, float b, CgShared<kThreads>& sh, int& parity) {
    a = cg_wave_sum(a);
    b = cg_wave_sum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh.red[parity][w][0] = a;
        sh.red[parity][w][1] = b;
    }
    __syncthreads();
    float2 t = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i) {
        t.x += sh.red[parity][i][0];
        t.y += sh.red[parity][i][1];
    }
    parity ^= 1;
    return t;
}

// Solves the coarsest system for right-hand side B (bricks) into X (bricks,
// DOF positions only); C.ndof <= kMaxDof (thread t owns rows t + j*kThreads,
// so a smaller kMaxDof only drops row slots that would be empty: the results
// do not depend on it). Every thread of the block must call it. kKeep: the
// solution, as X would hold it, also goes to sh.sp[d] (d < ndof; the zero slot
// stays zero), and X may be null; the caller's next barrier publishes it.
template <int kThreads, class V = float, int kMaxDof = kMaxCoarsestDof, bool kKeep = false>
__device__ void coarsest_cg_block(const CoarsestView& C, const V* B, V* X, CgShared<kThreads>& sh) {
    static_assert(kMaxDof <= kMaxCoarsestDof && kMaxDof % kThreads == 0, "row slots");
    constexpr int kPer = kMaxDof / kThreads;
    int parity = 0;
    const int n = C.ndof;
    const int tid = threadIdx.x;
    int nb[kPer][6];
    float cf[kPer][8], x[kPer], r[kPer], z[kPer], p[kPer], q[kPer];
    constexpr int kZero = kMaxCoarsestDof;
    if (tid == 0) sh.sp[kZero] = 0.f;   // published by the first reduction's barrier
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int d = tid + j * kThreads;
        x[j] = r[j] = z[j] = p[j] = q[j] = 0.f;
#pragma unroll
        for (int k = 0; k < 6; ++k) nb[j][k] = kZero;
#pragma unroll
        for (int k = 0; k < 8; ++k) cf[j][k] = 0.f;
        if (d < n) {
            // rows are 24 B (nb) and 32 B (coef) at 256-byte aligned bases
            const int2* nrow = reinterpret_cast<const int2*>(C.nb) + size_t(d) * 3;
            const float4* crow = reinterpret_cast<const float4*>(C.coef) + size_t(d) * 2;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int2 t = nrow[k];
                nb[j][2 * k] = t.x >= 0 ? t.x : kZero;
                nb[j][2 * k + 1] = t.y >= 0 ? t.y : kZero;
            }
            const float4 c0 = crow[0], c1 = crow[1];
            cf[j][0] = c0.x; cf[j][1] = c0.y; cf[j][2] = c0.z; cf[j][3] = c0.w;
            cf[j][4] = c1.x; cf[j][5] = c1.y; cf[j][6] = c1.z; cf[j][7] = c1.w;
            r[j] = float(B[C.pos[d]]);
        }
    }
    // ||b||^2 and <r, z> of z = M r in one reduction, whose barrier publishes z
    float rr = 0.f, rz = 0.f;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int d = tid + j * kThreads;
        if (d < n) {
            rr += r[j] * r[j];
            z[j] = cf[j][7] * r[j];
            sh.sp[d] = z[j];
            rz += r[j] * z[j];
        }
    }
    const float2 s0 = cg_sum2<kThreads>(rr, rz, sh, parity);
    const float rhs_norm2 = s0.x;
    const float threshold = fmaxf(C.tol * C.tol * rhs_norm2, FLT_MIN);
    if (rhs_norm2 != 0.f && !(rhs_norm2 < threshold)) {
        float abs_new = s0.y, beta = 0.f;
        for (int it = 0; it < C.max_iters; ++it) {
            // p = z + beta p and A p = A z + beta A p (beta = 0 on the first step)
            float pq = 0.f;
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int d = tid + j * kThreads;
                if (d < n) {
                    float v[6];
#pragma unroll
                    for (int k = 0; k < 6; ++k) v[k] = sh.sp[nb[j][k]];
                    const float f1 = v[0] * cf[j][1], f2 = v[1] * cf[j][2], f3 = v[2] * cf[j][3];
                    const float f4 = v[3] * cf[j][4], f5 = v[4] * cf[j][5], f6 = v[5] * cf[j][6];
                    const float s = ((f1 + f2) + (f3 + f4)) + (f5 + f6);
                    const float az = fmaf(z[j], cf[j][0], s);
                    p[j] = z[j] + beta * p[j];
                    q[j] = az + beta * q[j];
                    pq += p[j] * q[j];
                }
            }
            // the barrier also orders this step's reads of z before the writes below
            const float alpha = abs_new / cg_sum2<kThreads>(pq, 0.f, sh, parity).x;
            rr = 0.f;
            rz = 0.f;
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int d = tid + j * kThreads;
                if (d < n) {
                    x[j] = x[j] + alpha * p[j];
                    r[j] = r[j] - alpha * q[j];
                    rr += r[j] * r[j];
                    z[j] = cf[j][7] * r[j];
                    sh.sp[d] = z[j];
                    rz += r[j] * z[j];
                }
            }
            const float2 sums = cg_sum2<kThreads>(rr, rz, sh, parity);
            if (sums.x < threshold) break;
            const float abs_old = abs_new;
            abs_new = sums.y;
            beta = abs_new / abs_old;
        }
    }
    // every thread has passed the last barrier that precedes a read of sh.sp
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int d = tid + j * kThreads;
        if (d < n) {
            const V xv = V(x[j]);
            if constexpr (kKeep) {
                sh.sp[d] = float(xv);
                if (X) X[C.pos[d]] = xv;
            } else {
                X[C.pos[d]] = xv;
            }
        }
    }
}

}  // namespace dev
}  // namespace uaamg
