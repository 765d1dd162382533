// The bottom of the W-cycle in one launch.
//
// The two levels at the bottom of the hierarchy are small (the coarsest holds
// <= 4000 DOFs and the one above it roughly 8x that), so a separate launch per
// sweep there costs more in launch latency than in work. The tail kernel runs
// both mu = 2 visits of level L-2 (muCyclePreconditioner<2,...> called twice
// from level L-3, simd_vdb_poisson.cpp:1932-1935) on one 512-thread workgroup:
// pre-smoothing, residual + restriction, the coarsest solve, prolongation fused
// into the first post-smoothing sweep, two more sweeps — with a workgroup
// barrier between phases (512 threads: the coarsest rows of 8 DOFs per
// thread stay in registers). The arithmetic per value is exactly that of the
// per-sweep kernels (same device routines), only the launch structure changes.
#include "coarsest_device.h"
#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

constexpr int kTailThreads = 512;
constexpr int kTailWaves = kTailThreads / 64;

struct TailShared {
    float rs[kTailWaves][512];
    CgShared<kTailThreads> cg;
};

__device__ __forceinline__ int first_wave_leaf() { return __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class Src>
__device__ void block_jacobi(const LevelView& L, const Src& src, const float* B, float* Y, float w) {
    const int lane = threadIdx.x & 63;
    for (int leaf = first_wave_leaf(); leaf < L.nleaf; leaf += kTailWaves) {
        const uint32_t rm = row_mask(L, leaf, lane);
        if (rm) {
            const LeafRec R = leaf_rec(L, leaf);
            float xc[8], s[8], y[8];
            stencil_offdiag(L, R, src, leaf, lane, xc, s);
            jacobi_row(L, R, leaf, lane, xc, s, B, w, y);
            mask_row(rm, y);
            st8(Y + size_t(leaf) * 512 + lane * 8, y);
        }
    }
}

// r = b - A x per leaf, restricted through the wave's own LDS tile into the
// parent's octant (same order as the residual_restrict kernel).
__device__ void block_residual_restrict(const LevelView& F, const float* X, const float* B, float* Cv,
                                        float (*rs)[512]) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    for (int leaf = first_wave_leaf(); leaf < F.nleaf; leaf += kTailWaves) {
        float r[8];
        fill8(r, 0.f);
        const uint32_t rm = row_mask(F, leaf, lane);
        if (rm) {
            const LeafRec R = leaf_rec(F, leaf);
            float xc[8], s[8], dg[8], b[8];
            stencil_offdiag(F, R, PlainSrc{X}, leaf, lane, xc, s);
            ld8_slot(F.diag, R.sd, lane, F.dflt_diag, dg);
            ld8(B + size_t(leaf) * 512 + lane * 8, b);
#pragma unroll
            for (int z = 0; z < 8; ++z) r[z] = b[z] - fmaf(xc[z], dg[z], s[z]);
            mask_row(rm, r);
        }
        wave_sync();   // the previous leaf's reads of the tile are done
#pragma unroll
        for (int z = 0; z < 8; ++z) rs[w][lane * 8 + z] = r[z];
        wave_sync();
        const int cx = lane >> 4, cy = (lane >> 2) & 3, cz = lane & 3;
        float sum = 0.f;
#pragma unroll
        for (int ii = 0; ii < 2; ++ii)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int kk = 0; kk < 2; ++kk) sum += rs[w][((2 * cx + ii) * 8 + (2 * cy + jj)) * 8 + 2 * cz + kk];
        const int p = F.parent[leaf];
        if (p >= 0) {
            const int cl = p >> 3, oct = p & 7;
            const int off = ((cx + 4 * ((oct >> 2) & 1)) * 8 + (cy + 4 * ((oct >> 1) & 1))) * 8 + cz + 4 * (oct & 1);
            Cv[size_t(cl) * 512 + off] = sum * 0.125f;
        }
    }
}

__global__ __launch_bounds__(kTailThreads) void tail_kernel(TailArgs A) {
    __shared__ TailShared sh;
    const LevelView& L = A.L;
    float* cur = A.X;
    auto next = [&]() { return cur == A.T ? A.U : A.T; };
    for (int v = 0; v < A.nvisits; ++v) {
        if (v == 0 && A.first_skip) {
            float* dst = next();
            block_jacobi(L, FirstJacobiSrc{A.B, L.invdiag, L.dflt_invdiag, A.wpre[0]}, A.B, dst, A.wpre[1]);
            __syncthreads();
            cur = dst;
            dst = next();
            block_jacobi(L, PlainSrc{cur}, A.B, dst, A.wpre[2]);
            __syncthreads();
            cur = dst;
        } else {
            for (int i = 0; i < 3; ++i) {
                float* dst = next();
                block_jacobi(L, PlainSrc{cur}, A.B, dst, A.wpre[i]);
                __syncthreads();
                cur = dst;
            }
        }
        block_residual_restrict(L, cur, A.B, A.Bc, sh.rs);
        __syncthreads();
        coarsest_cg_block<kTailThreads>(A.C, A.Bc, A.Xc, sh.cg);
        __syncthreads();
        {
            float* dst = next();
            block_jacobi(L, ProlongSrc{cur, A.Xc, L.mask}, A.B, dst, A.wpost[0]);
            __syncthreads();
            cur = dst;
            dst = next();
            block_jacobi(L, PlainSrc{cur}, A.B, dst, A.wpost[1]);
            __syncthreads();
            cur = dst;
        }
        block_jacobi(L, PlainSrc{cur}, A.B, A.X, A.wpost[2]);
        __syncthreads();
        cur = A.X;
    }
}

}  // namespace

hipError_t launch_tail(const TailArgs& a, hipStream_t s) {
    if (a.L.nleaf <= 0) return hipSuccess;
    if (a.C.ndof > kMaxCoarsestDof) return hipErrorInvalidValue;
    tail_kernel<<<1, kTailThreads, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace uaamg
