// Variational viscosity kernels: the 15-point staggered operator in every
// applier mode, the transfers between levels, the coarsest Eigen-CG
// restatement. Arithmetic per value follows light_weight_applier::
// component_operator<c> (simd_viscosity3d.cpp:2841-3582): products rounded one
// by one, the same-channel sum ((x+ + x-) + (y+ + y-)) + (z+ + z-), then the two
// cross-channel groups, fused multiply-adds exactly where the AVX code uses
// _mm256_fmadd_ps (build with -ffp-contract=off).
#include <cfloat>

#include "viscosity_host.h"
#include "viscosity_kernels.h"
#include "wave_reduce.h"
#include "xcd_map.h"

namespace uaamg {
namespace {

constexpr int kVWaves = 4;
constexpr int kVThreads = 64 * kVWaves;

__device__ __forceinline__ int nb_slot(int ax, int ay, int az) { return (ax + 1) * 9 + (ay + 1) * 3 + (az + 1); }

// leaf holding leaf-local row (X, Y) in [-1, 8]^2, shifted by az leaves in z
__device__ __forceinline__ int row_leaf(const ViscLevelView& L, int leaf, int X, int Y, int az) {
    const int ax = X < 0 ? -1 : (X > 7 ? 1 : 0);
    const int ay = Y < 0 ? -1 : (Y > 7 ? 1 : 0);
    if (!ax && !ay && !az) return leaf;
    return L.nb27[size_t(leaf) * 27 + nb_slot(ax, ay, az)];
}

__device__ __forceinline__ int row_off(int X, int Y) { return ((X & 7) * 8 + (Y & 7)) * 8; }

__device__ __forceinline__ void ld8v(const float* p, float v[8]) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    const float4 b = *reinterpret_cast<const float4*>(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void st8v(float* p, const float v[8]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
}

// value source: a level vector channel (missing leaf = 0) or a coefficient
// grid (missing leaf or slot = the level default)
struct VecSrc {
    const float* v;
    int ch;
    __device__ __forceinline__ const float* base(const ViscLevelView& L, int lf) const {
        return (lf >= 0 && lf < L.nleaf_dof) ? v + (size_t(lf) * 3 + ch) * 512 : nullptr;
    }
    __device__ __forceinline__ float missing(const ViscLevelView&) const { return 0.f; }
};
template <int GID>
struct CoefSrc {
    __device__ __forceinline__ const float* base(const ViscLevelView& L, int lf) const {
        if (lf < 0) return nullptr;
        const int s = L.slot[size_t(lf) * 32 + GID];
        return s >= 0 ? L.pool + size_t(s) * 512 : nullptr;
    }
    __device__ __forceinline__ float missing(const ViscLevelView& L) const { return L.dflt[GID]; }
};

// the 8 values at (lx + OX, ly + OY, z + OZ), z = 0..7
template <int OX, int OY, int OZ, class Src>
__device__ __forceinline__ void fetch(const ViscLevelView& L, const Src& S, int leaf, int lx, int ly, float out[8]) {
    const int X = lx + OX, Y = ly + OY;
    const float* b = S.base(L, row_leaf(L, leaf, X, Y, 0));
    float row[8];
    if (b) ld8v(b + row_off(X, Y), row);
    else {
        const float m = S.missing(L);
#pragma unroll
        for (int k = 0; k < 8; ++k) row[k] = m;
    }
    if (OZ == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) out[k] = row[k];
    } else if (OZ > 0) {
#pragma unroll
        for (int k = 0; k < 7; ++k) out[k] = row[k + 1];
        const float* nb = S.base(L, row_leaf(L, leaf, X, Y, 1));
        out[7] = nb ? nb[row_off(X, Y)] : S.missing(L);
    } else {
#pragma unroll
        for (int k = 7; k > 0; --k) out[k] = row[k - 1];
        const float* nb = S.base(L, row_leaf(L, leaf, X, Y, -1));
        out[0] = nb ? nb[row_off(X, Y) + 7] : S.missing(L);
    }
}

template <int A>
struct Unit {
    static constexpr int x = A == 0, y = A == 1, z = A == 2;
};

// One row (lane) of the applier on channel C of `leaf`: Y = op(X, B).
template <int MODE, int C>
__device__ __forceinline__ void visc_row(const ViscLevelView& L, int leaf, int lane, const float* Xv, const float* Bv,
                                         float* Yv) {
    const int lx = lane >> 3, ly = lane & 7;
    const uint32_t rm = uint32_t(L.mask[(size_t(leaf) * 3 + C) * 8 + lx] >> (ly * 8)) & 0xffu;
    if (!rm) return;   // rows without a DOF are not written
    constexpr int d0 = C, d1 = (C + 1) % 3, d2 = (C + 2) % 3;
    using E0 = Unit<d0>;
    using E1 = Unit<d1>;
    using E2 = Unit<d2>;
    const VecSrc xc{Xv, C}, x1{Xv, d1}, x2{Xv, d2};

    float s[8];
    {
        float a[8], wa[8], b[8], wb[8];
        float f12[8], f34[8];
        fetch<1, 0, 0>(L, xc, leaf, lx, ly, a);
        fetch<1, 0, 0>(L, CoefSrc<kVNeg + 3 * C + 0>{}, leaf, lx, ly, wa);
        fetch<-1, 0, 0>(L, xc, leaf, lx, ly, b);
        fetch<0, 0, 0>(L, CoefSrc<kVNeg + 3 * C + 0>{}, leaf, lx, ly, wb);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            f12[k] = p + q;
        }
        fetch<0, 1, 0>(L, xc, leaf, lx, ly, a);
        fetch<0, 1, 0>(L, CoefSrc<kVNeg + 3 * C + 1>{}, leaf, lx, ly, wa);
        fetch<0, -1, 0>(L, xc, leaf, lx, ly, b);
        fetch<0, 0, 0>(L, CoefSrc<kVNeg + 3 * C + 1>{}, leaf, lx, ly, wb);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            f34[k] = p + q;
        }
        fetch<0, 0, 1>(L, xc, leaf, lx, ly, a);
        fetch<0, 0, 1>(L, CoefSrc<kVNeg + 3 * C + 2>{}, leaf, lx, ly, wa);
        fetch<0, 0, -1>(L, xc, leaf, lx, ly, b);
        fetch<0, 0, 0>(L, CoefSrc<kVNeg + 3 * C + 2>{}, leaf, lx, ly, wb);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            const float f56 = p + q;
            const float t = f12[k] + f34[k];
            s[k] = t + f56;
        }
    }
    {
        // normal d2: channel d1 at v, v + e1, v - e0, v - e0 + e1
        float u[8], c[8], t01[8];
        fetch<0, 0, 0>(L, x1, leaf, lx, ly, u);
        fetch<0, 0, 0>(L, CoefSrc<kVCross + 4 * d2 + 0>{}, leaf, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) t01[k] = u[k] * c[k];
        fetch<E1::x, E1::y, E1::z>(L, x1, leaf, lx, ly, u);
        fetch<0, 0, 0>(L, CoefSrc<kVCross + 4 * d2 + 1>{}, leaf, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = c[k] * u[k];
            t01[k] = t01[k] + p;
        }
        float t23[8];
        fetch<-E0::x, -E0::y, -E0::z>(L, x1, leaf, lx, ly, u);
        fetch<-E0::x, -E0::y, -E0::z>(L, CoefSrc<kVCross + 4 * d2 + 2>{}, leaf, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) t23[k] = c[k] * u[k];
        fetch<E1::x - E0::x, E1::y - E0::y, E1::z - E0::z>(L, x1, leaf, lx, ly, u);
        fetch<-E0::x, -E0::y, -E0::z>(L, CoefSrc<kVCross + 4 * d2 + 3>{}, leaf, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = c[k] * u[k];
            t23[k] = t23[k] + p;
            const float T2 = t01[k] + t23[k];
            s[k] = s[k] + T2;
        }
    }
    {
        // normal d1: channel d2 at v, v - e0, v + e2, v - e0 + e2
        float u[8], c[8], t01[8];
        fetch<0, 0, 0>(L, x2, leaf, lx, ly, u);
        fetch<0, 0, 0>(L, CoefSrc<kVCross + 4 * d1 + 0>{}, leaf, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) t01[k] = c[k] * u[k];
        fetch<-E0::x, -E0::y, -E0::z>(L, x2, leaf, lx, ly, u);
        fetch<-E0::x, -E0::y, -E0::z>(L, CoefSrc<kVCross + 4 * d1 + 1>{}, leaf, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = c[k] * u[k];
            t01[k] = t01[k] + p;
        }
        float t23[8];
        fetch<E2::x, E2::y, E2::z>(L, x2, leaf, lx, ly, u);
        fetch<0, 0, 0>(L, CoefSrc<kVCross + 4 * d1 + 2>{}, leaf, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) t23[k] = c[k] * u[k];
        fetch<E2::x - E0::x, E2::y - E0::y, E2::z - E0::z>(L, x2, leaf, lx, ly, u);
        fetch<-E0::x, -E0::y, -E0::z>(L, CoefSrc<kVCross + 4 * d1 + 3>{}, leaf, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = c[k] * u[k];
            t23[k] = t23[k] + p;
            const float T1 = t01[k] + t23[k];
            s[k] = s[k] + T1;
        }
    }

    const size_t row = (size_t(leaf) * 3 + C) * 512 + size_t(lx * 8 + ly) * 8;
    float lhs[8], y[8];
    ld8v(Xv + row, lhs);
    if (MODE == kVRedGS || MODE == kVBlackGS) {
        float b[8], inv[8];
        ld8v(Bv + row, b);
        fetch<0, 0, 0>(L, CoefSrc<kVInvDiag + C>{}, leaf, lx, ly, inv);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float t = b[k] - s[k];
            t = t * inv[k];
            t = t * 1.0f;
            const float r = fmaf(lhs[k], 0.0f, t);
            const bool red = ((lx + ly + k) & 1) == 0;
            y[k] = (red == (MODE == kVRedGS)) ? r : lhs[k];
        }
    } else {
        float dg[8];
        fetch<0, 0, 0>(L, CoefSrc<kVDiag + C>{}, leaf, lx, ly, dg);
        if (MODE == kVNormal) {
#pragma unroll
            for (int k = 0; k < 8; ++k) y[k] = fmaf(lhs[k], dg[k], s[k]);
        } else {
            float b[8];
            ld8v(Bv + row, b);
            if (MODE == kVResidual) {
#pragma unroll
                for (int k = 0; k < 8; ++k) y[k] = b[k] - fmaf(lhs[k], dg[k], s[k]);
            } else {
                float sc[8];
                if (MODE == kVJacobi) fetch<0, 0, 0>(L, CoefSrc<kVInvDiag + C>{}, leaf, lx, ly, sc);
                else fetch<0, 0, 0>(L, CoefSrc<kVSpai0 + C>{}, leaf, lx, ly, sc);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const float a = fmaf(lhs[k], dg[k], s[k]);
                    float t = b[k] - a;
                    t = t * sc[k];
                    y[k] = fmaf(1.0f, t, lhs[k]);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (!((rm >> k) & 1u)) y[k] = 0.f;
    st8v(Yv + row, y);
}

template <int MODE, int C>
__global__ __launch_bounds__(kVThreads) void visc_apply_kernel(ViscLevelView L, const float* Xv, const float* Bv,
                                                               float* Yv) {
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    if (w >= L.chan_off[C + 1] - L.chan_off[C]) return;
    visc_row<MODE, C>(L, L.chan_leaves[L.chan_off[C] + w], threadIdx.x & 63, Xv, Bv, Yv);
}

// restriction of one (coarse leaf, channel) row: 1/8 of the on fine children
__device__ __forceinline__ void visc_restrict_row(const ViscLevelView& Cl, const ViscLevelView& F, int leaf, int c,
                                                  int lane, const float* fine, float* coarse) {
    const int lx = lane >> 3, ly = lane & 7;
    const uint32_t rm = uint32_t(Cl.mask[(size_t(leaf) * 3 + c) * 8 + lx] >> (ly * 8)) & 0xffu;
    if (!rm) return;
    float out[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        out[z] = 0.f;
        if (!((rm >> z) & 1u)) continue;
        const int oct = ((lx >> 2) << 2) | ((ly >> 2) << 1) | (z >> 2);
        const int f = Cl.child[size_t(leaf) * 8 + oct];
        if (f < 0) continue;
        const uint64_t* fm = F.mask + (size_t(f) * 3 + c) * 8;
        const float* fv = fine + (size_t(f) * 3 + c) * 512;
        const int fx = 2 * (lx & 3), fy = 2 * (ly & 3), fz = 2 * (z & 3);
        float sum = 0.f;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    const int fo = (fx + a) * 64 + (fy + b) * 8 + fz + d;
                    if ((fm[fo >> 6] >> (fo & 63)) & 1u) sum += fv[fo];
                }
        out[z] = sum * 0.125f * 0.5f * 2.0f;
    }
    st8v(coarse + (size_t(leaf) * 3 + c) * 512 + size_t(lx * 8 + ly) * 8, out);
}

// prolongation<true> of one (fine leaf, channel) row
__device__ __forceinline__ void visc_prolong_row(const ViscLevelView& F, int leaf, int c, int lane, const float* coarse,
                                                 float* fine) {
    const int lx = lane >> 3, ly = lane & 7;
    const uint32_t rm = uint32_t(F.mask[(size_t(leaf) * 3 + c) * 8 + lx] >> (ly * 8)) & 0xffu;
    if (!rm) return;
    const int p = F.parent[leaf];
    if (p < 0) return;
    const int cl = p >> 3, oct = p & 7;
    const int cx = ((oct >> 2) & 1) * 4 + (lx >> 1), cy = ((oct >> 1) & 1) * 4 + (ly >> 1), cz0 = (oct & 1) * 4;
    const float* cv = coarse + (size_t(cl) * 3 + c) * 512 + cx * 64 + cy * 8 + cz0;
    float* row = fine + (size_t(leaf) * 3 + c) * 512 + size_t(lx * 8 + ly) * 8;
    float v[8];
    ld8v(row, v);
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if ((rm >> z) & 1u) v[z] += cv[z >> 1];
    st8v(row, v);
}

// ---------------------------------------------------- small-level visits
// On the coarse levels a W-cycle visit is ~29 dependent launches (12 colour
// passes, residual, restriction, prolongation, 12 colour passes). When every
// channel of a level has at most a few dozen leaves, one workgroup runs the
// whole half-visit with a workgroup barrier between passes: the same row
// routine, the same pass order, the same values.
constexpr int kSmallThreads = 512;
constexpr int kSmallWaves = kSmallThreads / 64;

template <int MODE, int C>
__device__ __forceinline__ void small_pass(const ViscLevelView& L, const float* X, const float* B, float* Y) {
    const int w0 = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = L.chan_off[C + 1] - L.chan_off[C];
    for (int w = w0; w < n; w += kSmallWaves) visc_row<MODE, C>(L, L.chan_leaves[L.chan_off[C] + w], lane, X, B, Y);
    __syncthreads();
}

__device__ __forceinline__ void small_rbgs(const ViscLevelView& L, float* X, const float* B, bool xyz) {
    if (xyz) {
        small_pass<kVRedGS, 0>(L, X, B, X);
        small_pass<kVBlackGS, 0>(L, X, B, X);
        small_pass<kVRedGS, 1>(L, X, B, X);
        small_pass<kVBlackGS, 1>(L, X, B, X);
        small_pass<kVRedGS, 2>(L, X, B, X);
        small_pass<kVBlackGS, 2>(L, X, B, X);
    } else {
        small_pass<kVBlackGS, 2>(L, X, B, X);
        small_pass<kVRedGS, 2>(L, X, B, X);
        small_pass<kVBlackGS, 1>(L, X, B, X);
        small_pass<kVRedGS, 1>(L, X, B, X);
        small_pass<kVBlackGS, 0>(L, X, B, X);
        small_pass<kVRedGS, 0>(L, X, B, X);
    }
}

// (zero X when precon,) 2 XYZ sweeps, residual into T, restriction into Cb
__global__ __launch_bounds__(kSmallThreads) void visc_small_pre_kernel(ViscLevelView L, ViscLevelView Cl, float* X,
                                                                       const float* B, float* T, float* Cb,
                                                                       int precon) {
    if (precon) {
        const size_t n4 = size_t(L.nleaf_dof) * 3 * 512 / 4;
        float4* x4 = reinterpret_cast<float4*>(X);
        for (size_t i = threadIdx.x; i < n4; i += kSmallThreads) x4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
    }
    small_rbgs(L, X, B, true);
    small_rbgs(L, X, B, true);
    small_pass<kVResidual, 0>(L, X, B, T);
    small_pass<kVResidual, 1>(L, X, B, T);
    small_pass<kVResidual, 2>(L, X, B, T);
    const int w0 = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int w = w0; w < Cl.nleaf_dof * 3; w += kSmallWaves) visc_restrict_row(Cl, L, w / 3, w % 3, lane, T, Cb);
}

// X += P Xc, then 2 ZYX sweeps
__global__ __launch_bounds__(kSmallThreads) void visc_small_post_kernel(ViscLevelView L, float* X, const float* B,
                                                                        const float* Xc) {
    const int w0 = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int w = w0; w < L.nleaf_dof * 3; w += kSmallWaves) visc_prolong_row(L, w / 3, w % 3, lane, Xc, X);
    __syncthreads();
    small_rbgs(L, X, B, false);
    small_rbgs(L, X, B, false);
}

// restriction: coarse (leaf, channel) waves
__global__ __launch_bounds__(kVThreads) void visc_restrict_kernel(ViscLevelView Cl, ViscLevelView F, const float* fine,
                                                                  float* coarse) {
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    if (w >= Cl.nleaf_dof * 3) return;
    visc_restrict_row(Cl, F, w / 3, w - 3 * (w / 3), threadIdx.x & 63, fine, coarse);
}

// prolongation<true>: fine (leaf, channel) waves add the parent coarse value
__global__ __launch_bounds__(kVThreads) void visc_prolong_kernel(ViscLevelView F, const float* coarse, float* fine) {
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    if (w >= F.nleaf_dof * 3) return;
    visc_prolong_row(F, w / 3, w - 3 * (w / 3), threadIdx.x & 63, coarse, fine);
}

__global__ __launch_bounds__(kVThreads) void visc_zero_kernel(float4* v, size_t n4) {
    const size_t i = size_t(blockIdx.x) * kVThreads + threadIdx.x;
    if (i < n4) v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------- coarsest solve
// Eigen::ConjugateGradient<RowMajor, Upper|Lower> with the diagonal
// preconditioner, zero guess, 10 iterations, tolerance = float epsilon
// (construct_exact_solver / solve_exact, simd_viscosity3d.cpp:2598-2657).
// Rows are summed in column order as Eigen's row-major product does; the
// dot-product order of Eigen's vectorised reductions is not reproduced.
constexpr int kCgT = 512;
constexpr int kCgPer = kViscMaxCoarsest / kCgT;

__device__ __forceinline__ float wsum(float v) { return dev::wave_sum_dpp(v); }

__device__ __forceinline__ float2 bsum2(float a, float b, float (*red)[kCgT / 64][2], int& par) {
    a = wsum(a);
    b = wsum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[par][w][0] = a;
        red[par][w][1] = b;
    }
    __syncthreads();
    float2 t = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < kCgT / 64; ++i) {
        t.x += red[par][i][0];
        t.y += red[par][i][1];
    }
    par ^= 1;
    return t;
}

__global__ __launch_bounds__(kCgT) void visc_coarsest_kernel(ViscCoarsestView C, const float* B, float* X) {
    __shared__ float sp[kViscMaxCoarsest];
    __shared__ float red[2][kCgT / 64][2];
    int par = 0;
    const int n = C.n, tid = threadIdx.x;
    float x[kCgPer], r[kCgPer], p[kCgPer], q[kCgPer], inv[kCgPer];
    float part = 0.f;
#pragma unroll
    for (int j = 0; j < kCgPer; ++j) {
        const int d = tid + j * kCgT;
        x[j] = r[j] = p[j] = q[j] = inv[j] = 0.f;
        if (d < n) {
            r[j] = B[C.pos[d]];
            inv[j] = C.invdiag[d];
            part += r[j] * r[j];
        }
    }
    const float rhs2 = bsum2(part, 0.f, red, par).x;
    const float thr = fmaxf(C.tol * C.tol * rhs2, FLT_MIN);
    if (rhs2 != 0.f && !(rhs2 < thr)) {
        part = 0.f;
#pragma unroll
        for (int j = 0; j < kCgPer; ++j) {
            const int d = tid + j * kCgT;
            if (d < n) {
                p[j] = inv[j] * r[j];
                sp[d] = p[j];
                part += r[j] * p[j];
            }
        }
        float abs_new = bsum2(part, 0.f, red, par).x;
        for (int it = 0; it < C.max_iters; ++it) {
            part = 0.f;
#pragma unroll
            for (int j = 0; j < kCgPer; ++j) {
                const int d = tid + j * kCgT;
                if (d < n) {
                    float t = 0.f;
                    for (int64_t k = C.ptr[d]; k < C.ptr[d + 1]; ++k) {
                        const float m = C.val[k] * sp[C.col[k]];
                        t = t + m;
                    }
                    q[j] = t;
                    part += p[j] * q[j];
                }
            }
            const float alpha = abs_new / bsum2(part, 0.f, red, par).x;
            float rr = 0.f, rz = 0.f;
#pragma unroll
            for (int j = 0; j < kCgPer; ++j) {
                const int d = tid + j * kCgT;
                if (d < n) {
                    x[j] = x[j] + alpha * p[j];
                    r[j] = r[j] - alpha * q[j];
                    rr += r[j] * r[j];
                    q[j] = inv[j] * r[j];
                    rz += r[j] * q[j];
                }
            }
            const float2 sums = bsum2(rr, rz, red, par);
            if (sums.x < thr) break;
            const float abs_old = abs_new;
            abs_new = sums.y;
            const float beta = abs_new / abs_old;
#pragma unroll
            for (int j = 0; j < kCgPer; ++j) {
                const int d = tid + j * kCgT;
                if (d < n) {
                    p[j] = q[j] + beta * p[j];
                    sp[d] = p[j];
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int j = 0; j < kCgPer; ++j) {
        const int d = tid + j * kCgT;
        if (d < n) X[C.pos[d]] = x[j];
    }
}

template <int MODE>
hipError_t apply_mode(const ViscLevelView& L, int c, const float* x, const float* b, float* y, hipStream_t s) {
    const int n = L.chan_off[c + 1] - L.chan_off[c];
    if (n <= 0) return hipSuccess;
    const unsigned g = unsigned((n + kVWaves - 1) / kVWaves);
    switch (c) {
        case 0: visc_apply_kernel<MODE, 0><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        case 1: visc_apply_kernel<MODE, 1><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        default: visc_apply_kernel<MODE, 2><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_visc_apply(const ViscLevelView& L, int mode, int c, const float* x, const float* b, float* y,
                             hipStream_t s) {
    if (c < 0 || c > 2) return hipErrorInvalidValue;
    switch (mode) {
        case kVNormal: return apply_mode<kVNormal>(L, c, x, b, y, s);
        case kVResidual: return apply_mode<kVResidual>(L, c, x, b, y, s);
        case kVJacobi: return apply_mode<kVJacobi>(L, c, x, b, y, s);
        case kVSpai0Mode: return apply_mode<kVSpai0Mode>(L, c, x, b, y, s);
        case kVRedGS: return apply_mode<kVRedGS>(L, c, x, b, y, s);
        case kVBlackGS: return apply_mode<kVBlackGS>(L, c, x, b, y, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_visc_restrict(const ViscLevelView& coarse, const ViscLevelView& fine, const float* fine_v,
                                float* coarse_v, hipStream_t s) {
    const int n = coarse.nleaf_dof * 3;
    if (n <= 0) return hipSuccess;
    visc_restrict_kernel<<<unsigned((n + kVWaves - 1) / kVWaves), kVThreads, 0, s>>>(coarse, fine, fine_v, coarse_v);
    return hipGetLastError();
}

hipError_t launch_visc_prolong(const ViscLevelView& fine, const ViscLevelView&, const float* coarse_v, float* fine_v,
                               hipStream_t s) {
    const int n = fine.nleaf_dof * 3;
    if (n <= 0) return hipSuccess;
    visc_prolong_kernel<<<unsigned((n + kVWaves - 1) / kVWaves), kVThreads, 0, s>>>(fine, coarse_v, fine_v);
    return hipGetLastError();
}

hipError_t launch_visc_coarsest(const ViscCoarsestView& C, const float* b, float* x, hipStream_t s) {
    if (C.n <= 0) return hipSuccess;
    if (C.n > kViscMaxCoarsest) return hipErrorInvalidValue;
    visc_coarsest_kernel<<<1, kCgT, 0, s>>>(C, b, x);
    return hipGetLastError();
}

hipError_t launch_visc_small_pre(const ViscLevelView& L, const ViscLevelView& coarse, float* x, const float* b,
                                 float* t, float* coarse_b, bool precon, hipStream_t s) {
    visc_small_pre_kernel<<<1, kSmallThreads, 0, s>>>(L, coarse, x, b, t, coarse_b, precon ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_visc_small_post(const ViscLevelView& L, float* x, const float* b, const float* xc, hipStream_t s) {
    visc_small_post_kernel<<<1, kSmallThreads, 0, s>>>(L, x, b, xc);
    return hipGetLastError();
}

hipError_t launch_visc_zero(const ViscLevelView& L, float* v, hipStream_t s) {
    const size_t n4 = size_t(L.nleaf_dof) * 3 * 512 / 4;
    if (!n4) return hipSuccess;
    visc_zero_kernel<<<unsigned((n4 + kVThreads - 1) / kVThreads), kVThreads, 0, s>>>(reinterpret_cast<float4*>(v),
                                                                                     n4);
    return hipGetLastError();
}

}  // namespace uaamg
