// Device kernels of the UAAMG-PCG Poisson solve (see poisson_kernels.h).
//
// Arithmetic follows LaplacianApplySIMD::LightWeightApplier
// (simd_vdb_poisson.cpp:755-1031) term by term: the six neighbour products are
// rounded separately and summed as ((x+ + x-) + (y+ + y-)) + (z+ + z-), then
// A*x = fma(x, diag, sum). The file is compiled with -ffp-contract=off so the
// only fused multiply-adds are the explicit fmaf calls.
#include <cfloat>

#include "poisson_kernels.h"

namespace uaamg {
namespace {

__device__ __forceinline__ int wave_leaf() {
    int leaf = int(blockIdx.x) * kWavesPerBlock + int(threadIdx.x >> 6);
    return __builtin_amdgcn_readfirstlane(leaf);
}

__device__ __forceinline__ void ld8(const float* p, float v[8]) {
    const float4 a = reinterpret_cast<const float4*>(p)[0];
    const float4 b = reinterpret_cast<const float4*>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

__device__ __forceinline__ void st8(float* p, const float v[8]) {
    reinterpret_cast<float4*>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<float4*>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}

__device__ __forceinline__ void fill8(float v[8], float c) {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = c;
}

__device__ __forceinline__ uint32_t row_mask(const LevelView& L, int leaf, int lane) {
    const uint64_t w = L.mask[size_t(leaf) * 8 + (lane >> 3)];
    return uint32_t(w >> ((lane & 7) * 8)) & 0xffu;
}

// load a row of a brick array, or a constant when the slot is absent
__device__ __forceinline__ void ld8_slot(const float* base, int slot, int row, float dflt, float v[8]) {
    if (slot >= 0) ld8(base + size_t(slot) * 512 + row * 8, v);
    else fill8(v, dflt);
}

// Own row of x and the off-diagonal sum of the 7-point stencil for one row.
__device__ __forceinline__ void stencil_offdiag(const LevelView& L, const float* X, int leaf, int lane, float xc[8],
                                                float s[8]) {
    const int lx = lane >> 3, ly = lane & 7;
    const int32_t* nb = L.nbr + size_t(leaf) * 6;
    const int nmx = nb[0], npx = nb[1], nmy = nb[2], npy = nb[3], nmz = nb[4], npz = nb[5];
    ld8(X + size_t(leaf) * 512 + lane * 8, xc);

    float xm[8], xp[8], ym[8], yp[8];
    {
        const int sl = lx > 0 ? leaf : nmx, sr = lx > 0 ? lane - 8 : lane + 56;
        if (sl >= 0) ld8(X + size_t(sl) * 512 + sr * 8, xm);
        else fill8(xm, 0.f);
    }
    {
        const int sl = lx < 7 ? leaf : npx, sr = lx < 7 ? lane + 8 : lane - 56;
        if (sl >= 0) ld8(X + size_t(sl) * 512 + sr * 8, xp);
        else fill8(xp, 0.f);
    }
    {
        const int sl = ly > 0 ? leaf : nmy, sr = ly > 0 ? lane - 1 : lane + 7;
        if (sl >= 0) ld8(X + size_t(sl) * 512 + sr * 8, ym);
        else fill8(ym, 0.f);
    }
    {
        const int sl = ly < 7 ? leaf : npy, sr = ly < 7 ? lane + 1 : lane - 7;
        if (sl >= 0) ld8(X + size_t(sl) * 512 + sr * 8, yp);
        else fill8(yp, 0.f);
    }
    const float zlo = nmz >= 0 ? X[size_t(nmz) * 512 + lane * 8 + 7] : 0.f;
    const float zhi = npz >= 0 ? X[size_t(npz) * 512 + lane * 8] : 0.f;

    const int32_t* sl = L.slot + size_t(leaf) * kSlotStride;
    const int sx = sl[kSlotX], sy = sl[kSlotY], sz = sl[kSlotZ];
    const int sxp = npx >= 0 ? L.slot[size_t(npx) * kSlotStride + kSlotX] : -1;
    const int syp = npy >= 0 ? L.slot[size_t(npy) * kSlotStride + kSlotY] : -1;
    const int szp = npz >= 0 ? L.slot[size_t(npz) * kSlotStride + kSlotZ] : -1;
    const float d = L.dflt_off;

    float cxm[8], cxp[8], cym[8], cyp[8], czm[8];
    ld8_slot(L.cx, sx, lane, d, cxm);
    ld8_slot(L.cx, lx < 7 ? sx : sxp, lx < 7 ? lane + 8 : lane - 56, d, cxp);
    ld8_slot(L.cy, sy, lane, d, cym);
    ld8_slot(L.cy, ly < 7 ? sy : syp, ly < 7 ? lane + 1 : lane - 7, d, cyp);
    ld8_slot(L.cz, sz, lane, d, czm);
    const float czhi = szp >= 0 ? L.cz[size_t(szp) * 512 + lane * 8] : d;

#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const float zpv = z < 7 ? xc[z + 1] : zhi;
        const float czp = z < 7 ? czm[z + 1] : czhi;
        const float zmv = z > 0 ? xc[z - 1] : zlo;
        const float f1 = xp[z] * cxp[z];
        const float f2 = xm[z] * cxm[z];
        const float f3 = yp[z] * cyp[z];
        const float f4 = ym[z] * cym[z];
        const float f5 = zpv * czp;
        const float f6 = zmv * czm[z];
        s[z] = ((f1 + f2) + (f3 + f4)) + (f5 + f6);
    }
}

template <int MODE, bool kLevel0>
__global__ __launch_bounds__(kBlockThreads) void apply_kernel(LevelView L, const float* X, const float* B, float* Y,
                                                              float w, float one_minus_w) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    if (rm == 0) return;
    float xc[8], s[8], y[8];
    stencil_offdiag(L, X, leaf, lane, xc, s);
    const int sd = L.slot[size_t(leaf) * kSlotStride + kSlotDiag];
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
                ld8_slot(L.spai0, L.slot[size_t(leaf) * kSlotStride + kSlotSpai0], lane, L.dflt_spai0, sp);
#pragma unroll
                for (int z = 0; z < 8; ++z) {
                    const float r = b[z] - ax[z];
                    y[z] = fmaf(sp[z], r, xc[z]);
                }
            }
        }
    }
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
    if (rm == 0) return;
    const size_t row = size_t(leaf) * 512 + lane * 8;
    float b[8], c[8], y[8];
    ld8(B + row, b);
    const int32_t* sl = L.slot + size_t(leaf) * kSlotStride;
    if (kind == 1) {
        ld8_slot(L.spai0, sl[kSlotSpai0], lane, L.dflt_spai0, c);
#pragma unroll
        for (int z = 0; z < 8; ++z) y[z] = c[z] * b[z];
    } else {
        ld8_slot(L.invdiag, sl[kSlotDiag], lane, L.dflt_invdiag, c);
        const int par = ((lane >> 3) + (lane & 7)) & 1;
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            float v = c[z] * b[z];
            v = w * v;
            if (kind == 2 && ((par + z) & 1)) v = 0.f;
            y[z] = v;
        }
    }
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
    if (rm == 0) return;
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
    if (rm == 0) return;
    const int p = Fl.parent[leaf];
    if (p < 0) return;
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

__device__ __forceinline__ int64_t row_dof_base(const LevelView& L, int leaf, int lane) {
    const uint64_t* m = L.mask + size_t(leaf) * 8;
    const int lx = lane >> 3, ly = lane & 7;
    int64_t base = L.dofbase[leaf];
    for (int w = 0; w < lx; ++w) base += __popcll(m[w]);
    const uint64_t below = ly == 0 ? 0ull : (m[lx] & ((uint64_t(1) << (ly * 8)) - 1));
    return base + __popcll(below);
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

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Sum over the block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[kWavesPerBlock];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kWavesPerBlock; ++i) t += part[i];
    return t;
}

// Publishes this block's partial; returns true in every thread of the block
// that completes the grid. The last block reads all partials after an
// agent-scope acquire fence, so the final sum has a fixed order.
__device__ __forceinline__ bool publish_partial(double* partials, unsigned int* counter, double block_value) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = block_value;
        __threadfence();
        const unsigned int prev = atomicAdd(counter, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    return last;
}

// Fixed-order sum of n partials by the whole block; valid in thread 0.
__device__ __forceinline__ double sum_partials(const double* partials, int n) {
    __threadfence();
    double t = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) t += __builtin_nontemporal_load(partials + i);
    __syncthreads();
    return block_sum(t);
}

template <int TARGET>
__global__ __launch_bounds__(kBlockThreads) void dot_kernel(LevelView L, const float* A, const float* Bv,
                                                            PcgScalars* sc, double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    double acc = 0;
    if (leaf < L.nleaf) {
        const uint32_t rm = row_mask(L, leaf, lane);
        if (rm) {
            float a[8], b[8];
            const size_t row = size_t(leaf) * 512 + lane * 8;
            ld8(A + row, a);
            ld8(Bv + row, b);
#pragma unroll
            for (int z = 0; z < 8; ++z)
                if ((rm >> z) & 1u) acc += double(a[z]) * double(b[z]);
        }
    }
    const double bs = block_sum(acc);
    if (publish_partial(partials, &sc->counter[TARGET], bs)) {
        const double total = sum_partials(partials, gridDim.x);
        if (threadIdx.x == 0) {
            if (TARGET == kDotRho) {
                sc->rho = total;
            } else {
                const float rho_new = float(total), rho = float(sc->rho);
                sc->beta = rho_new / rho;
                sc->rho_new = total;
                sc->rho = total;
            }
            sc->counter[TARGET] = 0;
        }
    }
}

__global__ __launch_bounds__(kBlockThreads) void laplacian_dot_kernel(LevelView L, const float* P, float* Z,
                                                                      PcgScalars* sc, double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    double acc = 0;
    if (leaf < L.nleaf) {
        const uint32_t rm = row_mask(L, leaf, lane);
        if (rm) {
            float xc[8], s[8], dg[8], y[8];
            stencil_offdiag(L, P, leaf, lane, xc, s);
            ld8_slot(L.diag, L.slot[size_t(leaf) * kSlotStride + kSlotDiag], lane, L.dflt_diag, dg);
#pragma unroll
            for (int z = 0; z < 8; ++z) {
                const bool on = (rm >> z) & 1u;
                y[z] = on ? fmaf(xc[z], dg[z], s[z]) : 0.f;
                if (on) acc += double(xc[z]) * double(y[z]);
            }
            st8(Z + size_t(leaf) * 512 + lane * 8, y);
        }
    }
    const double bs = block_sum(acc);
    if (publish_partial(partials, &sc->counter[2], bs)) {
        const double total = sum_partials(partials, gridDim.x);
        if (threadIdx.x == 0) {
            sc->sigma = total;
            sc->alpha = float(sc->rho) / float(total);
            sc->counter[2] = 0;
        }
    }
}

// max|r| with NaN/Inf propagation (grid_abs_max_op, openvdb_grid_math_op.h:4-40)
template <bool AXPY>
__global__ __launch_bounds__(kBlockThreads) void absmax_kernel(LevelView L, const float* Zv, float* R,
                                                               PcgScalars* sc, double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    float m = 0.f;
    float bad = 0.f;
    if (leaf < L.nleaf) {
        const uint32_t rm = row_mask(L, leaf, lane);
        if (rm) {
            const size_t row = size_t(leaf) * 512 + lane * 8;
            float r[8];
            ld8(R + row, r);
            if (AXPY) {
                const float na = -sc->alpha;
                float z[8];
                ld8(Zv + row, z);
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if ((rm >> k) & 1u) {
                        const float t = na * z[k];
                        r[k] = r[k] + t;
                    }
                st8(R + row, r);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if ((rm >> k) & 1u) {
                    const float a = fabsf(r[k]);
                    if (!(a <= FLT_MAX)) bad = 1.f;
                    else m = fmaxf(m, a);
                }
        }
    }
    m = wave_max(m);
    bad = wave_max(bad);
    // pack into a double: partial = bad ? -1 : m
    __shared__ float wm[kWavesPerBlock];
    if (lane == 0) wm[threadIdx.x >> 6] = bad > 0.f ? -1.f : m;
    __syncthreads();
    double bv = 0;
    if (threadIdx.x == 0) {
        float v = 0.f;
        bool b = false;
        for (int i = 0; i < kWavesPerBlock; ++i) {
            if (wm[i] < 0.f) b = true;
            else v = fmaxf(v, wm[i]);
        }
        bv = b ? -1.0 : double(v);
    }
    if (publish_partial(partials, &sc->counter[3], bv)) {
        __threadfence();
        float v = 0.f;
        bool b = false;
        for (int i = threadIdx.x; i < int(gridDim.x); i += blockDim.x) {
            const double p = __builtin_nontemporal_load(partials + i);
            if (p < 0) b = true;
            else v = fmaxf(v, float(p));
        }
        __shared__ float tv[kBlockThreads];
        __shared__ int tb[kBlockThreads];
        tv[threadIdx.x] = v;
        tb[threadIdx.x] = b;
        __syncthreads();
        if (threadIdx.x == 0) {
            float mv = 0.f;
            bool any = false;
            for (int i = 0; i < kBlockThreads; ++i) {
                mv = fmaxf(mv, tv[i]);
                any = any || tb[i];
            }
            sc->nu = any ? __builtin_nanf("") : mv;
            sc->counter[3] = 0;
        }
    }
}

template <bool UPDATE_P>
__global__ __launch_bounds__(kBlockThreads) void update_kernel(LevelView L, const float* Zv, float* X, float* P,
                                                               const PcgScalars* sc) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    if (!rm) return;
    const float alpha = sc->alpha;
    const size_t row = size_t(leaf) * 512 + lane * 8;
    float x[8], p[8];
    ld8(X + row, x);
    ld8(P + row, p);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((rm >> k) & 1u) {
            const float t = alpha * p[k];
            x[k] = x[k] + t;
        }
    st8(X + row, x);
    if (UPDATE_P) {
        const float beta = sc->beta;
        float z[8];
        ld8(Zv + row, z);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if ((rm >> k) & 1u) {
                const float t = beta * p[k];
                p[k] = z[k] + t;
            }
        st8(P + row, p);
    }
}

// ---------------------------------------------------------- coarsest solve
// Eigen::ConjugateGradient with DiagonalPreconditioner, zero initial guess,
// maxIterations = 10, tolerance = float epsilon, restated
// (constructCoarsestLevelExactSolver, simd_vdb_poisson.cpp:2114-2126).
constexpr int kCgThreads = 1024;

__device__ __forceinline__ float cg_block_sum(float v, float* red) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = 0.f;
#pragma unroll
    for (int i = 0; i < kCgThreads / 64; ++i) t += red[i];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kCgThreads) void coarsest_cg_kernel(CoarsestView C, const float* B, float* X) {
    __shared__ float sx[kMaxCoarsestDof], sr[kMaxCoarsestDof], sp[kMaxCoarsestDof], sz[kMaxCoarsestDof];
    __shared__ float red[kCgThreads / 64];
    const int n = C.ndof;
    const int tid = threadIdx.x;
    float part = 0.f;
    for (int d = tid; d < n; d += kCgThreads) {
        const float b = B[C.pos[d]];
        sr[d] = b;
        sx[d] = 0.f;
        part += b * b;
    }
    __syncthreads();
    const float rhs_norm2 = cg_block_sum(part, red);
    bool done = rhs_norm2 == 0.f;
    const float threshold = fmaxf(C.tol * C.tol * rhs_norm2, FLT_MIN);
    if (!done && rhs_norm2 < threshold) done = true;
    if (!done) {
        part = 0.f;
        for (int d = tid; d < n; d += kCgThreads) {
            const float p = C.coef[size_t(d) * 8 + 7] * sr[d];
            sp[d] = p;
            part += sr[d] * p;
        }
        __syncthreads();
        float abs_new = cg_block_sum(part, red);
        for (int it = 0; it < C.max_iters; ++it) {
            part = 0.f;
            for (int d = tid; d < n; d += kCgThreads) {
                const int32_t* nb = C.nb + size_t(d) * 6;
                const float* cf = C.coef + size_t(d) * 8;
                float v[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) v[k] = nb[k] >= 0 ? sp[nb[k]] : 0.f;
                const float f1 = v[0] * cf[1], f2 = v[1] * cf[2], f3 = v[2] * cf[3], f4 = v[3] * cf[4];
                const float f5 = v[4] * cf[5], f6 = v[5] * cf[6];
                const float s = ((f1 + f2) + (f3 + f4)) + (f5 + f6);
                const float t = fmaf(sp[d], cf[0], s);
                sz[d] = t;
                part += sp[d] * t;
            }
            __syncthreads();
            const float alpha = abs_new / cg_block_sum(part, red);
            part = 0.f;
            for (int d = tid; d < n; d += kCgThreads) {
                sx[d] = sx[d] + alpha * sp[d];
                const float r = sr[d] - alpha * sz[d];
                sr[d] = r;
                part += r * r;
            }
            __syncthreads();
            const float res_norm2 = cg_block_sum(part, red);
            if (res_norm2 < threshold) break;
            part = 0.f;
            for (int d = tid; d < n; d += kCgThreads) {
                const float z = C.coef[size_t(d) * 8 + 7] * sr[d];
                sz[d] = z;
                part += sr[d] * z;
            }
            __syncthreads();
            const float abs_old = abs_new;
            abs_new = cg_block_sum(part, red);
            const float beta = abs_new / abs_old;
            for (int d = tid; d < n; d += kCgThreads) sp[d] = sz[d] + beta * sp[d];
            __syncthreads();
        }
    }
    for (int d = tid; d < n; d += kCgThreads) X[C.pos[d]] = sx[d];
}

inline unsigned grid_for(int32_t nleaf) { return unsigned((nleaf + kWavesPerBlock - 1) / kWavesPerBlock); }

}  // namespace

int reduction_blocks(const LevelView& L) { return int(grid_for(L.nleaf)) > 0 ? int(grid_for(L.nleaf)) : 1; }

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

hipError_t launch_coarsest_cg(const CoarsestView& C, const float* b, float* x, hipStream_t s) {
    if (C.ndof <= 0) return hipSuccess;
    if (C.ndof > kMaxCoarsestDof) return hipErrorInvalidValue;
    coarsest_cg_kernel<<<1, kCgThreads, 0, s>>>(C, b, x);
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

hipError_t launch_dot(const LevelView& L, const float* a, const float* b, PcgScalars* sc, double* partials,
                      int target, hipStream_t s) {
    const unsigned g = unsigned(reduction_blocks(L));
    if (target == kDotRho) dot_kernel<kDotRho><<<g, kBlockThreads, 0, s>>>(L, a, b, sc, partials);
    else dot_kernel<kDotRhoNew><<<g, kBlockThreads, 0, s>>>(L, a, b, sc, partials);
    return hipGetLastError();
}

hipError_t launch_laplacian_dot(const LevelView& L, const float* p, float* z, PcgScalars* sc, double* partials,
                                hipStream_t s) {
    laplacian_dot_kernel<<<unsigned(reduction_blocks(L)), kBlockThreads, 0, s>>>(L, p, z, sc, partials);
    return hipGetLastError();
}

hipError_t launch_axpy_absmax(const LevelView& L, const float* z, float* r, PcgScalars* sc, double* partials,
                              hipStream_t s) {
    absmax_kernel<true><<<unsigned(reduction_blocks(L)), kBlockThreads, 0, s>>>(L, z, r, sc, partials);
    return hipGetLastError();
}

hipError_t launch_absmax(const LevelView& L, const float* r, PcgScalars* sc, double* partials, hipStream_t s) {
    absmax_kernel<false><<<unsigned(reduction_blocks(L)), kBlockThreads, 0, s>>>(L, nullptr, const_cast<float*>(r),
                                                                                 sc, partials);
    return hipGetLastError();
}

hipError_t launch_update_xp(const LevelView& L, const float* z, float* x, float* p, const PcgScalars* sc,
                            hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    update_kernel<true><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, z, x, p, sc);
    return hipGetLastError();
}

hipError_t launch_update_x(const LevelView& L, const float* p, float* x, const PcgScalars* sc, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    update_kernel<false><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, nullptr, x, const_cast<float*>(p), sc);
    return hipGetLastError();
}

}  // namespace uaamg
