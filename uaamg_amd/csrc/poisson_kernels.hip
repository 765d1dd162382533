// Device kernels of the UAAMG-PCG Poisson solve (see poisson_kernels.h).
//
// Arithmetic follows LaplacianApplySIMD::LightWeightApplier
// (simd_vdb_poisson.cpp:755-1031) term by term: the six neighbour products are
// rounded separately and summed as ((x+ + x-) + (y+ + y-)) + (z+ + z-), then
// A*x = fma(x, diag, sum). The file is compiled with -ffp-contract=off so the
// only fused multiply-adds are the explicit fmaf calls.
#include <cfloat>

#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

template <int MODE, bool kLevel0, class VX = float, class VB = float, class VY = float>
__global__ __launch_bounds__(kBlockThreads) void apply_kernel(LevelView L, const VX* X, const VB* B, VY* Y, float w,
                                                              float one_minus_w) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    // every load is issued unconditionally (all in bounds) and rows without a
    // DOF skip only the store, so no load waits for the mask
    const uint32_t rm = row_mask(L, leaf, lane);
    const LeafRec R = leaf_rec(L, leaf);
    float xc[8], s[8], y[8];
    stencil_offdiag(L, R, PlainSrcT<VX>{X}, leaf, lane, xc, s);
    const int sd = R.sd;
    const size_t row = size_t(leaf) * 512 + lane * 8;
    if (MODE == kModeRedGS || MODE == kModeBlackGS) {
        float b[8], inv[8];
        ld8(B + row, b);
        ld8_slot(L.invdiag, sd, lane, L.dflt_invdiag, inv);
        const int lx = lane >> 3, ly = lane & 7;
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            float r = b[z] - s[z];
            r = r * inv[z];
            r = r * w;
            const float upd = fmaf(xc[z], one_minus_w, r);
            const bool red = ((lx + ly + z) & 1) == 0;
            y[z] = (red == (MODE == kModeRedGS)) ? upd : xc[z];
        }
    } else {
        float dg[8];
        ld8_slot(L.diag, sd, lane, L.dflt_diag, dg);
        float ax[8];
#pragma unroll
        for (int z = 0; z < 8; ++z) ax[z] = fmaf(xc[z], dg[z], s[z]);
        if (MODE == kModeLaplacian) {
#pragma unroll
            for (int z = 0; z < 8; ++z) y[z] = ax[z];
        } else {
            float b[8];
            ld8(B + row, b);
            if (MODE == kModeResidual) {
#pragma unroll
                for (int z = 0; z < 8; ++z) y[z] = b[z] - ax[z];
            } else if (MODE == kModeJacobi) {
                float inv[8];
                ld8_slot(L.invdiag, sd, lane, L.dflt_invdiag, inv);
#pragma unroll
                for (int z = 0; z < 8; ++z) {
                    float r = b[z] - ax[z];
                    r = r * inv[z];
                    y[z] = fmaf(w, r, xc[z]);
                }
            } else {   // SPAI0
                float sp[8];
                ld8_slot(L.spai0, R.ssp, lane, L.dflt_spai0, sp);
#pragma unroll
                for (int z = 0; z < 8; ++z) {
                    const float r = b[z] - ax[z];
                    y[z] = fmaf(sp[z], r, xc[z]);
                }
            }
        }
    }
    if (rm == 0) return;
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if (!((rm >> z) & 1u)) y[z] = 0.f;
    st8(Y + row, y);
}

// kind: 0 Jacobi, 1 SPAI0, 2 red half of RBGS
__global__ __launch_bounds__(kBlockThreads) void first_sweep_kernel(LevelView L, int kind, const float* B, float* Y,
                                                                    float w) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    const LeafRec R = leaf_rec(L, leaf);
    const size_t row = size_t(leaf) * 512 + lane * 8;
    float b[8], c[8], y[8];
    ld8(B + row, b);
    if (kind == 1) {
        ld8_slot(L.spai0, R.ssp, lane, L.dflt_spai0, c);
#pragma unroll
        for (int z = 0; z < 8; ++z) y[z] = c[z] * b[z];
    } else {
        ld8_slot(L.invdiag, R.sd, lane, L.dflt_invdiag, c);
        const int par = ((lane >> 3) + (lane & 7)) & 1;
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            float v = c[z] * b[z];
            v = w * v;
            if (kind == 2 && ((par + z) & 1)) v = 0.f;
            y[z] = v;
        }
    }
    if (rm == 0) return;
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if (!((rm >> z) & 1u)) y[z] = 0.f;
    st8(Y + row, y);
}

__global__ __launch_bounds__(kBlockThreads) void restrict_kernel(LevelView C, const float* F, float* Cv) {
    const int leaf = wave_leaf();
    if (leaf >= C.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(C, leaf, lane);
    const int cx = lane >> 3, cy = lane & 7;
    float out[8];
#pragma unroll
    for (int bz = 0; bz < 2; ++bz) {
        const int oct = ((cx >= 4) << 2) | ((cy >= 4) << 1) | bz;
        const int fl = C.child[size_t(leaf) * 8 + oct];
        if (fl < 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) out[4 * bz + c] = 0.f;
            continue;
        }
        const int fx = 2 * (cx & 3), fy = 2 * (cy & 3);
        float v[2][2][8];
#pragma unroll
        for (int ii = 0; ii < 2; ++ii)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) ld8(F + size_t(fl) * 512 + ((fx + ii) * 8 + (fy + jj)) * 8, v[ii][jj]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float sum = 0.f;
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                    for (int kk = 0; kk < 2; ++kk) sum += v[ii][jj][2 * c + kk];
            out[4 * bz + c] = sum * 0.125f;
        }
    }
    if (rm == 0) return;
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if (!((rm >> z) & 1u)) out[z] = 0.f;
    st8(Cv + size_t(leaf) * 512 + lane * 8, out);
}

__global__ __launch_bounds__(kBlockThreads) void prolong_add_kernel(LevelView Fl, const float* Cv, float* Fv,
                                                                    float alpha) {
    const int leaf = wave_leaf();
    if (leaf >= Fl.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(Fl, leaf, lane);
    const int p = Fl.parent[leaf];
    if (p < 0 || rm == 0) return;
    const int cl = p >> 3, oct = p & 7;
    const int bx = (oct >> 2) & 1, by = (oct >> 1) & 1, bz = oct & 1;
    const int crow = (((lane >> 3) >> 1) + 4 * bx) * 8 + (((lane & 7) >> 1) + 4 * by);
    const float4 c4 = *reinterpret_cast<const float4*>(Cv + size_t(cl) * 512 + crow * 8 + 4 * bz);
    const float c[4] = {alpha * c4.x, alpha * c4.y, alpha * c4.z, alpha * c4.w};
    const size_t row = size_t(leaf) * 512 + lane * 8;
    float y[8];
    ld8(Fv + row, y);
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if ((rm >> z) & 1u) y[z] += c[z >> 1];
    st8(Fv + row, y);
}

// First two scheduled/weighted Jacobi sweeps from a zero guess in one pass:
// the first sweep's result is rebuilt where the second sweep reads it.
template <bool kLevel0, class VB = float, class VY = float>
__global__ __launch_bounds__(kBlockThreads) void jacobi_first2_kernel(LevelView L, const VB* B, VY* Y, float w0,
                                                                      float w1) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    const LeafRec R = leaf_rec(L, leaf);
    const FirstJacobiSrcT<VB> src{B, L.invdiag, L.dflt_invdiag, w0};
    float xc[8], s[8], y[8];
    stencil_offdiag(L, R, src, leaf, lane, xc, s);
    jacobi_row(L, R, leaf, lane, xc, s, B, w1, y);
    if (rm == 0) return;
    mask_row(rm, y);
    st8(Y + size_t(leaf) * 512 + lane * 8, y);
}

// Prolongation (x += P x_coarse) fused into the first post-smoothing sweep.
template <bool kLevel0, class V = float, class VB = float>
__global__ __launch_bounds__(kBlockThreads) void prolong_jacobi_kernel(LevelView L, const V* X, const V* Xc,
                                                                       const VB* B, V* Y, float w) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    const LeafRec R = leaf_rec(L, leaf);
    const ProlongSrcT<V> src{X, Xc, L.mask};
    float xc[8], s[8], y[8];
    stencil_offdiag(L, R, src, leaf, lane, xc, s);
    jacobi_row(L, R, leaf, lane, xc, s, B, w, y);
    if (rm == 0) return;
    mask_row(rm, y);
    st8(Y + size_t(leaf) * 512 + lane * 8, y);
}

// Residual r = b - A x of a fine leaf, restricted in the same pass: each
// coarse voxel's eight children lie in one fine leaf, so the leaf writes the
// 4^3 octant of its parent coarse leaf (restriction, :1624-1684, sums the
// children in x,y,z nested order and scales by 1/8).
template <bool kLevel0, class V = float, class VB = float>
__global__ __launch_bounds__(kBlockThreads) void residual_restrict_kernel(LevelView F, const V* X, const VB* B,
                                                                          V* Cv) {
    __shared__ float rs[kWavesPerBlock][512];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int leaf = wave_leaf();
    const bool valid = leaf < F.nleaf;
    float r[8];
    fill8(r, 0.f);
    int par = -1;
    if (valid) {
        const uint32_t rm = row_mask(F, leaf, lane);
        const LeafRec R = leaf_rec(F, leaf);
        par = R.par;
        float xc[8], s[8], dg[8], b[8];
        stencil_offdiag(F, R, PlainSrcT<V>{X}, leaf, lane, xc, s);
        ld8_slot(F.diag, R.sd, lane, F.dflt_diag, dg);
        ld8(B + size_t(leaf) * 512 + lane * 8, b);
#pragma unroll
        for (int z = 0; z < 8; ++z) r[z] = b[z] - fmaf(xc[z], dg[z], s[z]);
        mask_row(rm, r);
    }
#pragma unroll
    for (int z = 0; z < 8; ++z) rs[wave][lane * 8 + z] = r[z];
    __syncthreads();
    if (!valid) return;
    const int cx = lane >> 4, cy = (lane >> 2) & 3, cz = lane & 3;
    float sum = 0.f;
#pragma unroll
    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) sum += rs[wave][((2 * cx + ii) * 8 + (2 * cy + jj)) * 8 + 2 * cz + kk];
    const int p = par;
    if (p < 0) return;   // every DOF leaf has a parent; kept as a bounds guard
    const int cl = p >> 3, oct = p & 7;
    const int off = ((cx + 4 * ((oct >> 2) & 1)) * 8 + (cy + 4 * ((oct >> 1) & 1))) * 8 + cz + 4 * (oct & 1);
    st1(Cv + size_t(cl) * 512 + off, sum * 0.125f);
}


__global__ __launch_bounds__(kBlockThreads) void scatter_dof_kernel(LevelView L, const float* dv, float* Y) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    float y[8];
    fill8(y, 0.f);
    if (rm) {
        int64_t k = row_dof_base(L, leaf, lane);
#pragma unroll
        for (int z = 0; z < 8; ++z)
            if ((rm >> z) & 1u) y[z] = dv[k++];
    }
    st8(Y + size_t(leaf) * 512 + lane * 8, y);
}

__global__ __launch_bounds__(kBlockThreads) void gather_dof_kernel(LevelView L, const float* X, float* dv) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    if (!rm) return;
    float x[8];
    ld8(X + size_t(leaf) * 512 + lane * 8, x);
    int64_t k = row_dof_base(L, leaf, lane);
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if ((rm >> z) & 1u) dv[k++] = x[z];
}


}  // namespace

hipError_t launch_apply(const LevelView& L, int mode, const float* x, const float* b, float* y, float w,
                        hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    const dim3 g(grid_for(L.nleaf)), t(kBlockThreads);
    const float omw = 1.0f - w;
    if (L.level == 0) {
        switch (mode) {
            case kModeLaplacian: apply_kernel<kModeLaplacian, true><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeResidual: apply_kernel<kModeResidual, true><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeJacobi: apply_kernel<kModeJacobi, true><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeSpai0: apply_kernel<kModeSpai0, true><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeRedGS: apply_kernel<kModeRedGS, true><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeBlackGS: apply_kernel<kModeBlackGS, true><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (mode) {
            case kModeLaplacian: apply_kernel<kModeLaplacian, false><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeResidual: apply_kernel<kModeResidual, false><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeJacobi: apply_kernel<kModeJacobi, false><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeSpai0: apply_kernel<kModeSpai0, false><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeRedGS: apply_kernel<kModeRedGS, false><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            case kModeBlackGS: apply_kernel<kModeBlackGS, false><<<g, t, 0, s>>>(L, x, b, y, w, omw); break;
            default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

hipError_t launch_first_sweep(const LevelView& L, int kind, const float* b, float* y, float w, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    first_sweep_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, kind, b, y, w);
    return hipGetLastError();
}

hipError_t launch_restrict(const LevelView& coarse, const float* fine_v, float* coarse_v, hipStream_t s) {
    if (coarse.nleaf <= 0) return hipSuccess;
    restrict_kernel<<<grid_for(coarse.nleaf), kBlockThreads, 0, s>>>(coarse, fine_v, coarse_v);
    return hipGetLastError();
}

hipError_t launch_prolong_add(const LevelView& fine, const float* coarse_v, float* fine_v, float alpha,
                              hipStream_t s) {
    if (fine.nleaf <= 0) return hipSuccess;
    prolong_add_kernel<<<grid_for(fine.nleaf), kBlockThreads, 0, s>>>(fine, coarse_v, fine_v, alpha);
    return hipGetLastError();
}

hipError_t launch_scatter_dof(const LevelView& L, const float* dofvec, float* bricks, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    scatter_dof_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, dofvec, bricks);
    return hipGetLastError();
}

hipError_t launch_gather_dof(const LevelView& L, const float* bricks, float* dofvec, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    gather_dof_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, bricks, dofvec);
    return hipGetLastError();
}

hipError_t launch_jacobi_first2(const LevelView& L, const float* b, float* y, float w0, float w1, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    if (L.level == 0) jacobi_first2_kernel<true><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, b, y, w0, w1);
    else jacobi_first2_kernel<false><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, b, y, w0, w1);
    return hipGetLastError();
}

hipError_t launch_prolong_jacobi(const LevelView& L, const float* x, const float* xc, const float* b, float* y,
                                 float w, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    if (L.level == 0) prolong_jacobi_kernel<true><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, x, xc, b, y, w);
    else prolong_jacobi_kernel<false><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, x, xc, b, y, w);
    return hipGetLastError();
}

hipError_t launch_residual_restrict(const LevelView& fine, const float* x, const float* b, float* coarse_b,
                                    hipStream_t s) {
    if (fine.nleaf <= 0) return hipSuccess;
    if (fine.level == 0)
        residual_restrict_kernel<true><<<grid_for(fine.nleaf), kBlockThreads, 0, s>>>(fine, x, b, coarse_b);
    else residual_restrict_kernel<false><<<grid_for(fine.nleaf), kBlockThreads, 0, s>>>(fine, x, b, coarse_b);
    return hipGetLastError();
}

// ---- mixed-precision W-cycle launchers (level 0 reads the FP32 PCG residual)
hipError_t launch_jacobi_first2(const LevelView& L, const float* b, bf16* y, float w0, float w1, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    jacobi_first2_kernel<true, float, bf16><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, b, y, w0, w1);
    return hipGetLastError();
}

hipError_t launch_jacobi_first2(const LevelView& L, const bf16* b, bf16* y, float w0, float w1, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    jacobi_first2_kernel<false, bf16, bf16><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, b, y, w0, w1);
    return hipGetLastError();
}

template <bool kLevel0, class VX, class VB, class VY>
static hipError_t jacobi_t(const LevelView& L, const VX* x, const VB* b, VY* y, float w, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    apply_kernel<kModeJacobi, kLevel0, VX, VB, VY>
        <<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, x, b, y, w, 1.0f - w);
    return hipGetLastError();
}

hipError_t launch_jacobi(const LevelView& L, const bf16* x, const float* b, bf16* y, float w, hipStream_t s) {
    return jacobi_t<true>(L, x, b, y, w, s);
}
hipError_t launch_jacobi(const LevelView& L, const bf16* x, const float* b, float* y, float w, hipStream_t s) {
    return jacobi_t<true>(L, x, b, y, w, s);
}
hipError_t launch_jacobi(const LevelView& L, const float* x, const float* b, bf16* y, float w, hipStream_t s) {
    return jacobi_t<true>(L, x, b, y, w, s);
}
hipError_t launch_jacobi(const LevelView& L, const bf16* x, const bf16* b, bf16* y, float w, hipStream_t s) {
    return jacobi_t<false>(L, x, b, y, w, s);
}

hipError_t launch_residual_restrict(const LevelView& fine, const bf16* x, const float* b, bf16* coarse_b,
                                    hipStream_t s) {
    if (fine.nleaf <= 0) return hipSuccess;
    residual_restrict_kernel<true, bf16, float>
        <<<grid_for(fine.nleaf), kBlockThreads, 0, s>>>(fine, x, b, coarse_b);
    return hipGetLastError();
}

hipError_t launch_residual_restrict(const LevelView& fine, const bf16* x, const bf16* b, bf16* coarse_b,
                                    hipStream_t s) {
    if (fine.nleaf <= 0) return hipSuccess;
    residual_restrict_kernel<false, bf16, bf16>
        <<<grid_for(fine.nleaf), kBlockThreads, 0, s>>>(fine, x, b, coarse_b);
    return hipGetLastError();
}

hipError_t launch_prolong_jacobi(const LevelView& L, const bf16* x, const bf16* xc, const float* b, bf16* y, float w,
                                 hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    prolong_jacobi_kernel<true, bf16, float><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, x, xc, b, y, w);
    return hipGetLastError();
}

hipError_t launch_prolong_jacobi(const LevelView& L, const bf16* x, const bf16* xc, const bf16* b, bf16* y, float w,
                                 hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    prolong_jacobi_kernel<false, bf16, bf16><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, x, xc, b, y, w);
    return hipGetLastError();
}

}  // namespace uaamg
