// Outer-PCG kernels (dot products, ||r||_inf, vector updates, the fused
// PCG-boundary kernels) and the coarsest-level solve.
//
// Reductions are two launches: every block of the first writes one partial
// (no fences or atomics — on a multi-XCD part an agent-scope release per block
// writes back the L2 each time), then a one-block finalize kernel sums the
// partials in a fixed order and updates the scalars. The kernel boundary makes
// the partials visible once, and the fixed order makes every result
// reproducible run to run. Dot products accumulate in FP64.
#include <cfloat>
#include <cstdlib>
#include <cstring>

#include "batch_kernels.h"
#include "coarsest_device.h"
#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

enum FinalizeOp { kFinRho = 0, kFinRhoNew = 1, kFinSigma = 2, kFinNu = 3 };

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Block sum in a fixed order (wave butterflies, then the waves in order);
// the result is valid in thread 0.
template <int kThreads>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[kThreads / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kThreads / 64; ++i) t += part[i];
    return t;
}

// partial of the block's leaf range (indexed by the logical block, so the
// finalize sums in leaf order whatever the dispatch order)
__device__ __forceinline__ void write_partial(double* partials, double v) {
    if (threadIdx.x == 0) partials[logical_block()] = v;
}

template <int TARGET>
__global__ __launch_bounds__(kBlockThreads) void dot_kernel(LevelView L, const float* A, const float* Bv,
                                                            double* partials) {
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
    write_partial(partials, block_sum<kBlockThreads>(acc));
}

// z = A p and the partials of <p, z>
__global__ __launch_bounds__(kBlockThreads) void laplacian_dot_kernel(LevelView L, const float* P, float* Z,
                                                                      double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    double acc = 0;
    if (leaf < L.nleaf) {
        const uint32_t rm = row_mask(L, leaf, lane);
        const LeafRec R = leaf_rec(L, leaf);
        float xc[8], s[8], dg[8], y[8];
        stencil_offdiag(L, R, PlainSrc{P}, leaf, lane, xc, s);
        ld8_slot(L.diag, R.sd, lane, L.dflt_diag, dg);
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            const bool on = (rm >> z) & 1u;
            y[z] = on ? fmaf(xc[z], dg[z], s[z]) : 0.f;
            if (on) acc += double(xc[z]) * double(y[z]);
        }
        if (rm) st8(Z + size_t(leaf) * 512 + lane * 8, y);
    }
    write_partial(partials, block_sum<kBlockThreads>(acc));
}

// Last level-0 smoothing sweep of the W-cycle with the partials of <z, r>:
// z is the sweep's output and r its right-hand side.
template <class VX = float>
__global__ __launch_bounds__(kBlockThreads) void jacobi_dot_kernel(LevelView L, const VX* X, const float* B, float* Y,
                                                                   float w, double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    double acc = 0;
    if (leaf < L.nleaf) {
        const uint32_t rm = row_mask(L, leaf, lane);
        const LeafRec R = leaf_rec(L, leaf);
        float xc[8], s[8], y[8], b[8];
        stencil_offdiag(L, R, PlainSrcT<VX>{X}, leaf, lane, xc, s);
        jacobi_row(L, R, leaf, lane, xc, s, B, w, y);
        mask_row(rm, y);
        if (rm) st8(Y + size_t(leaf) * 512 + lane * 8, y);
        ld8(B + size_t(leaf) * 512 + lane * 8, b);
#pragma unroll
        for (int z = 0; z < 8; ++z)
            if ((rm >> z) & 1u) acc += double(y[z]) * double(b[z]);
    }
    write_partial(partials, block_sum<kBlockThreads>(acc));
}

// One PCG step across the iteration boundary: x += alpha p; p' = z + beta p;
// q = A p' and the partials of <p', q>. p' goes to a second buffer because
// neighbouring leaves still read p. At most 128 VGPRs, so four waves per
// SIMD fit (the default allocation takes 130: three waves).
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(4))) void update_laplacian_dot_kernel(
    LevelView L, const float* Z, const float* P, float* Pn, float* X, float* Q, const PcgScalars* sc,
    double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    double acc = 0;
    if (leaf < L.nleaf) {
        const uint32_t rm = row_mask(L, leaf, lane);
        const LeafRec R = leaf_rec(L, leaf);
        const float alpha = sc->alpha, beta = sc->beta;
        const size_t row = size_t(leaf) * 512 + lane * 8;
        float x[8], p[8];
        ld8(X + row, x);
        ld8(P + row, p);
        float pc[8], s[8], dg[8], q[8];
        stencil_offdiag(L, R, UpdateSrc{Z, P, beta}, leaf, lane, pc, s);
        ld8_slot(L.diag, R.sd, lane, L.dflt_diag, dg);
        mask_row(rm, pc);
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            const bool on = (rm >> z) & 1u;
            if (on) {
                const float t = alpha * p[z];
                x[z] = x[z] + t;
            }
            q[z] = on ? fmaf(pc[z], dg[z], s[z]) : 0.f;
            if (on) acc += double(pc[z]) * double(q[z]);
        }
        if (rm) {
            st8(X + row, x);
            st8(Pn + row, pc);
            st8(Q + row, q);
        }
    }
    write_partial(partials, block_sum<kBlockThreads>(acc));
}

// running max|r| of one lane; a NaN or Inf sets `bad`
__device__ __forceinline__ void absmax_acc(float r, float& m, float& bad) {
    const float a = fabsf(r);
    if (!(a <= FLT_MAX)) bad = 1.f;
    else m = fmaxf(m, a);
}

// the block's partial max|r|, -1 when any entry was not finite
__device__ __forceinline__ void write_absmax_partial(float m, float bad, double* partials) {
    m = wave_max(m);
    bad = wave_max(bad);
    __shared__ float wm[kWavesPerBlock];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = bad > 0.f ? -1.f : m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float v = 0.f;
        bool b = false;
        for (int i = 0; i < kWavesPerBlock; ++i) {
            if (wm[i] < 0.f) b = true;
            else v = fmaxf(v, wm[i]);
        }
        partials[logical_block()] = b ? -1.0 : double(v);
    }
}

// (optional r -= alpha z, then) partial max|r| with NaN/Inf flagged as -1
// (grid_abs_max_op, openvdb_grid_math_op.h:4-40)
template <bool AXPY>
__global__ __launch_bounds__(kBlockThreads) void absmax_kernel(LevelView L, const float* Zv, float* R,
                                                               const PcgScalars* sc, double* partials) {
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
                if ((rm >> k) & 1u) absmax_acc(r[k], m, bad);
        }
    }
    write_absmax_partial(m, bad, partials);
}

constexpr int kFinThreads = 256;

// Thread t's share of the partials: t, t + kFinThreads, ... For a sum the
// additions run in that order (the fixed order every reduction here uses);
// the loads are issued kFinAhead at a time so that they overlap, which does
// not change any sum. Level 0 at 512^3 has about 24500 partials: one load at
// a time took 35 us per reduction.
constexpr int kFinAhead = 8;

__device__ __forceinline__ double fin_sum(const double* part, int n) {
    double t = 0;
    int i = threadIdx.x;
    for (; i + (kFinAhead - 1) * kFinThreads < n; i += kFinAhead * kFinThreads) {
        double p[kFinAhead];
#pragma unroll
        for (int k = 0; k < kFinAhead; ++k) p[k] = part[i + k * kFinThreads];
#pragma unroll
        for (int k = 0; k < kFinAhead; ++k) t += p[k];
    }
    for (; i < n; i += kFinThreads) t += part[i];
    return t;
}

// The largest partial maximum (v) and whether any partial flags a NaN / Inf (b).
__device__ __forceinline__ void fin_max(const double* part, int n, float& v, int& b) {
    v = 0.f;
    b = 0;
    auto take = [&](double p) {
        if (p < 0) b = 1;
        else v = fmaxf(v, float(p));
    };
    int i = threadIdx.x;
    for (; i + (kFinAhead - 1) * kFinThreads < n; i += kFinAhead * kFinThreads) {
        double p[kFinAhead];
#pragma unroll
        for (int k = 0; k < kFinAhead; ++k) p[k] = part[i + k * kFinThreads];
#pragma unroll
        for (int k = 0; k < kFinAhead; ++k) take(p[k]);
    }
    for (; i < n; i += kFinThreads) take(part[i]);
}

// Fixed-order reduction of n partials and the scalar update of `op`:
// rho / (rho_new, beta = rho_new / rho, rho) / (sigma, alpha = rho / sigma) /
// nu, the last also published to the host mailbox as one 64-bit store.
__global__ __launch_bounds__(kFinThreads) void finalize_kernel(int op, int n, const double* partials,
                                                               PcgScalars* sc, HostMailbox* mb) {
    if (op == kFinNu) {
        __shared__ float tv[kFinThreads];
        __shared__ int tb[kFinThreads];
        float v;
        int b;
        fin_max(partials, n, v, b);
        tv[threadIdx.x] = v;
        tb[threadIdx.x] = b;
        __syncthreads();
        if (threadIdx.x == 0) {
            float mv = 0.f;
            int any = 0;
            for (int i = 0; i < kFinThreads; ++i) {
                mv = fmaxf(mv, tv[i]);
                any |= tb[i];
            }
            const float nu = any ? __builtin_nanf("") : mv;
            const unsigned int seq = sc->seq + 1;
            sc->nu = nu;
            sc->seq = seq;
            if (mb) {
                uint32_t bits;
                memcpy(&bits, &nu, 4);
                mb->word = (uint64_t(seq) << 32) | bits;
            }
        }
        return;
    }
    const double total = block_sum<kFinThreads>(fin_sum(partials, n));
    if (threadIdx.x != 0) return;
    switch (op) {
        case kFinRho: sc->rho = total; break;
        case kFinRhoNew: {
            const float rho_new = float(total), rho = float(sc->rho);
            sc->beta = rho_new / rho;
            sc->rho_new = total;
            sc->rho = total;
            break;
        }
        case kFinSigma:
            sc->sigma = total;
            sc->alpha = float(sc->rho) / float(total);
            break;
        default: break;
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
// Workgroup size of the coarsest CG: 512 threads by default; UAAMG_CG_THREADS = 256 or
// 1024 selects another (a tuning knob: the reductions' partial sums then group
// differently, so single and batched solves must run with the same setting).
int cg_threads() {
    static const int t = [] {
        const char* e = std::getenv("UAAMG_CG_THREADS");
        const int v = e ? std::atoi(e) : 512;
        return (v == 256 || v == 1024) ? v : 512;
    }();
    return t;
}

template <class V, int kMaxDof, int kT>
__global__ __launch_bounds__(kT) void coarsest_cg_kernel(CoarsestView C, const V* B, V* X) {
    __shared__ CgShared<kT> sh;
    coarsest_cg_block<kT, V, kMaxDof>(C, B, X, sh);
}

// r = b - A x with FP64 products and sums (fma(x, diag, s) of the operator
// becomes x * diag + s in FP64), rounded once to FP32, and the partial
// maxima of |r| as absmax_kernel forms them.
template <class Src>
__device__ __forceinline__ void residual64_row(const LevelView& L, const Src& src, const float* B, float* R, int leaf,
                                               int lane, uint32_t rm, float& m, float& bad) {
    const LeafRec Rec = leaf_rec(L, leaf);
    float r[8], xc[8], dg[8], b[8];
    double s[8];
    stencil_offdiag64(L, Rec, src, leaf, lane, xc, s);
    ld8_slot(L.diag, Rec.sd, lane, L.dflt_diag, dg);
    ld8(B + size_t(leaf) * 512 + lane * 8, b);
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const double ax = double(xc[z]) * double(dg[z]) + s[z];
        r[z] = float(double(b[z]) - ax);
    }
    mask_row(rm, r);
    st8(R + size_t(leaf) * 512 + lane * 8, r);
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if ((rm >> z) & 1u) absmax_acc(r[z], m, bad);
}

template <class Src>
__device__ __forceinline__ void residual64_body(const LevelView& L, const Src& src, const float* B, float* R,
                                                double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    float m = 0.f, bad = 0.f;
    if (leaf < L.nleaf) residual64_row(L, src, B, R, leaf, lane, row_mask(L, leaf, lane), m, bad);
    write_absmax_partial(m, bad, partials);
}

__global__ __launch_bounds__(kBlockThreads) void residual64_kernel(LevelView L, const float* X, const float* B,
                                                                   float* R, double* partials) {
    residual64_body(L, PlainSrc{X}, B, R, partials);
}

// The residual replacement inside a PCG recurrence: r = b - A (x + alpha p)
// as residual64_kernel forms it, for the iterate whose update x += alpha p is
// still pending (the next step or the final update stores exactly these
// values). Takes the place of r -= alpha A p.
__global__ __launch_bounds__(kBlockThreads) void replace64_kernel(LevelView L, const float* X, const float* P,
                                                                  const float* B, float* R, const PcgScalars* sc,
                                                                  double* partials) {
    residual64_body(L, StepSrc{X, P, sc->alpha}, B, R, partials);
}

// ---------------------------------------------------------- batched solve
// The same arithmetic as the single-problem kernels, with each leaf's scalars
// taken from its problem and the problem's state deciding what is updated.
__device__ __forceinline__ int leaf_problem(const int32_t* leafprob, int leaf) {
    return __builtin_amdgcn_readfirstlane(leafprob[leaf]);
}

__global__ __launch_bounds__(kBlockThreads) void bupdate_laplacian_dot_kernel(LevelView L, const int32_t* leafprob,
                                                                              const BatchPcg* st, const float* Z,
                                                                              const float* P, float* Pn, float* X,
                                                                              float* Q, double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    double acc = 0;
    if (leaf < L.nleaf) {
        const uint32_t rm = row_mask(L, leaf, lane);
        const LeafRec R = leaf_rec(L, leaf);
        const BatchPcg& sp = st[leaf_problem(leafprob, leaf)];
        const int state = sp.state;
        if (state == kBActive || state == kBConverging) {
            const float alpha = sp.alpha, beta = sp.beta;
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
            if (rm) st8(X + row, x);
            if (state == kBActive) {
                float pc[8], s[8], dg[8], q[8];
                stencil_offdiag(L, R, UpdateSrc{Z, P, beta}, leaf, lane, pc, s);
                mask_row(rm, pc);
                ld8_slot(L.diag, R.sd, lane, L.dflt_diag, dg);
#pragma unroll
                for (int z = 0; z < 8; ++z) {
                    const bool on = (rm >> z) & 1u;
                    q[z] = on ? fmaf(pc[z], dg[z], s[z]) : 0.f;
                    if (on) acc += double(pc[z]) * double(q[z]);
                }
                if (rm) {
                    st8(Pn + row, pc);
                    st8(Q + row, q);
                }
            }
        }
    }
    write_partial(partials, block_sum<kBlockThreads>(acc));
}

__global__ __launch_bounds__(kBlockThreads) void baxpy_absmax_kernel(LevelView L, const int32_t* leafprob,
                                                                     const BatchPcg* st, const float* Zv,
                                                                     const float* X, const float* P, const float* B,
                                                                     float* R, double* partials) {
    const int leaf = wave_leaf();
    const int lane = threadIdx.x & 63;
    float m = 0.f, bad = 0.f;
    if (leaf < L.nleaf) {
        const uint32_t rm = row_mask(L, leaf, lane);
        const BatchPcg& sp = st[leaf_problem(leafprob, leaf)];
        if (sp.state == kBActive && sp.replacing) {
            residual64_row(L, StepSrc{X, P, sp.alpha}, B, R, leaf, lane, rm, m, bad);
        } else if (rm && sp.state == kBActive) {
            const size_t row = size_t(leaf) * 512 + lane * 8;
            float r[8], z[8];
            ld8(R + row, r);
            const float na = -sp.alpha;
            ld8(Zv + row, z);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if ((rm >> k) & 1u) {
                    const float t = na * z[k];
                    r[k] = r[k] + t;
                }
            st8(R + row, r);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if ((rm >> k) & 1u) absmax_acc(r[k], m, bad);
        }
    }
    write_absmax_partial(m, bad, partials);
}

__global__ __launch_bounds__(kBlockThreads) void bupdate_x_kernel(LevelView L, const int32_t* leafprob,
                                                                  const BatchPcg* st, const float* P, float* X) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    const BatchPcg& sp = st[leaf_problem(leafprob, leaf)];
    if (!rm || sp.state != kBConverging) return;
    const float alpha = sp.alpha;
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
}

__global__ __launch_bounds__(kBlockThreads) void bupdate_xp_kernel(LevelView L, const int32_t* leafprob,
                                                                   const BatchPcg* st, const float* Zv, float* X,
                                                                   float* P) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_mask(L, leaf, lane);
    const BatchPcg& sp = st[leaf_problem(leafprob, leaf)];
    if (!rm || sp.state != kBActive) return;
    const float alpha = sp.alpha, beta = sp.beta;
    const size_t row = size_t(leaf) * 512 + lane * 8;
    float x[8], p[8], z[8];
    ld8(X + row, x);
    ld8(P + row, p);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((rm >> k) & 1u) {
            const float t = alpha * p[k];
            x[k] = x[k] + t;
        }
    st8(X + row, x);
    ld8(Zv + row, z);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((rm >> k) & 1u) {
            const float t = beta * p[k];
            p[k] = z[k] + t;
        }
    st8(P + row, p);
}

// One workgroup per problem. The sums and the scalar updates repeat
// finalize_kernel's on the problem's own block range; the state rules repeat
// the host loop of solve_pcg_dev / pcg_iterate.
__global__ __launch_bounds__(kFinThreads) void bfinalize_kernel(int op, const int32_t* blk_begin,
                                                                const double* partials, BatchPcg* st) {
    const int prob = blockIdx.x;
    BatchPcg& sp = st[prob];
    const int b0 = blk_begin[prob], n = blk_begin[prob + 1] - b0;
    const double* part = partials + b0;
    if (op == kBFinNu0 || op == kBFinNuFirst || op == kBFinNuIter || op == kBFinReplace || op == kBFinNuViscFirst ||
        op == kBFinNuViscIter) {
        __shared__ float tv[kFinThreads];
        __shared__ int tb[kFinThreads];
        float v;
        int b;
        fin_max(part, n, v, b);
        tv[threadIdx.x] = v;
        tb[threadIdx.x] = b;
        __syncthreads();
        if (threadIdx.x != 0) return;
        float mv = 0.f;
        int any = 0;
        for (int i = 0; i < kFinThreads; ++i) {
            mv = fmaxf(mv, tv[i]);
            any |= tb[i];
        }
        const float nu = any ? __builtin_nanf("") : mv;
        if (op == kBFinNu0) {
            sp.nu = sp.nu_prev = nu;
            sp.numax = sp.tol * nu;
            sp.iters = 0;
            sp.it0 = 0;
            sp.state = nu <= sp.numax ? kBSuccess : (sp.max_iters > 0 ? kBActive : kBFailed);
            return;
        }
        if (op == kBFinReplace) {   // the restart rule of solve_pcg_dev's residual replacement
            if (sp.state != kBSuccess || !sp.refine) return;
            sp.true_nu = nu;
            if (nu <= sp.numax || !(nu < 0.5f * sp.prev_true) || sp.restarts >= kBMaxReplacements ||
                sp.iters >= sp.max_iters) {
                sp.refine = 0;
                return;
            }
            sp.prev_true = nu;
            sp.restarts += 1;
            sp.replacements += 1;
            sp.it0 = sp.iters;
            sp.nu = sp.nu_prev = nu;
            sp.replacing = 0;
            sp.stalled = 0;
            sp.state = kBActive;
            return;
        }
        if (sp.state != kBActive) return;
        if (op == kBFinNuViscFirst || op == kBFinNuViscIter) {
            // the host loop of ViscDevice::solve_pcg_dev
            if (op == kBFinNuViscIter) sp.iters += 1;
            sp.nu = nu;
            if (nu <= sp.numax || sp.iters == sp.max_iters - 1) sp.state = kBConverging;
            return;
        }
        sp.nu = nu;
        if (op == kBFinNuIter) {
            sp.iters += 1;
            if (sp.iters >= sp.max_iters) {
                sp.state = kBFailed;
                return;
            }
        }
        // the step rules of pcg_iterate, the in-recurrence replacement (first
        // recurrence only) included
        const int replaced = sp.replacing;
        sp.replacing = 0;
        if (nu <= sp.numax) {
            sp.state = kBConverging;
        } else {
            if (replaced && !(nu < 0.5f * sp.nu_prev)) {
                // the true residual stopped halving: the recurrence ends for
                // a restart when the restart rule takes one (kBFinReplace
                // then finds this nu again), else goes on with the recursive residual
                if (sp.refine && sp.restarts < kBMaxReplacements && nu < 0.5f * sp.prev_true)
                    sp.state = kBConverging;
                else
                    sp.stalled = 1;
            } else if (nu > sp.nu_prev && sp.iters - sp.it0 > 3) {
                sp.state = kBFailed;
            }
            if (sp.state == kBActive && sp.refine && sp.restarts == 0 && !sp.stalled &&
                nu <= kReplaceBand * sp.numax) {
                sp.replacing = 1;
                sp.replacements += 1;
            }
        }
        sp.nu_prev = nu;
        return;
    }
    if (op == kBFinSigma && sp.state == kBConverging) {   // its last x update ran in this piece
        if (threadIdx.x == 0) sp.state = kBSuccess;
        return;
    }
    if (sp.state != kBActive) return;
    const double total = block_sum<kFinThreads>(fin_sum(part, n));
    if (threadIdx.x != 0) return;
    switch (op) {
        case kBFinRho: sp.rho = total; break;
        case kBFinRhoNew: {
            const float rho_new = float(total), rho = float(sp.rho);
            sp.beta = rho_new / rho;
            sp.rho_new = total;
            sp.rho = total;
            break;
        }
        case kBFinSigma:
            sp.sigma = total;
            sp.alpha = float(sp.rho) / float(total);
            break;
        default: break;
    }
}

// converging -> success after the final x update
__global__ __launch_bounds__(kFinThreads) void bcommit_kernel(int nprob, BatchPcg* st) {
    for (int p = threadIdx.x; p < nprob; p += kFinThreads)
        if (st[p].state == kBConverging) st[p].state = kBSuccess;
}

__global__ __launch_bounds__(kFinThreads) void bstatus_kernel(int nprob, const BatchPcg* st, unsigned int* seq,
                                                              HostMailbox* mb) {
    __shared__ int cnt[kFinThreads];
    int c = 0;
    for (int p = threadIdx.x; p < nprob; p += kFinThreads) c += st[p].state == kBActive ? 1 : 0;
    cnt[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x != 0) return;
    int total = 0;
    for (int i = 0; i < kFinThreads; ++i) total += cnt[i];
    const unsigned int q = *seq + 1;
    *seq = q;
    mb->word = (uint64_t(q) << 32) | uint32_t(total);
}

template <class V, int kMaxDof, int kT>
__global__ __launch_bounds__(kT) void bcoarsest_kernel(const CoarsestView* views, const V* B, V* X) {
    __shared__ CgShared<kT> sh;
    const CoarsestView C = views[blockIdx.x];
    coarsest_cg_block<kT, V, kMaxDof>(C, B, X, sh);
}

// one workgroup per system (grid n), the widest system ndof rows; the 4096-row
// form keeps at least 512 threads (16 rows a thread would not fit the registers)
template <class V>
hipError_t launch_cg_variant(bool batched, const CoarsestView* views, const CoarsestView& C, unsigned n, int ndof,
                             const V* b, V* x, hipStream_t s) {
    const int t = cg_threads();
    const bool small = ndof <= kCgSmallDof;
#define UAAMG_CG_LAUNCH(MAXDOF, T)                                                                 \
    do {                                                                                          \
        if (batched) bcoarsest_kernel<V, MAXDOF, T><<<n, T, 0, s>>>(views, b, x);                  \
        else coarsest_cg_kernel<V, MAXDOF, T><<<1, T, 0, s>>>(C, b, x);                            \
    } while (0)
    if (small) {
        if (t == 256) UAAMG_CG_LAUNCH(kCgSmallDof, 256);
        else if (t == 1024) UAAMG_CG_LAUNCH(kCgSmallDof, 1024);
        else UAAMG_CG_LAUNCH(kCgSmallDof, 512);
    } else {
        if (t == 1024) UAAMG_CG_LAUNCH(kMaxCoarsestDof, 1024);
        else UAAMG_CG_LAUNCH(kMaxCoarsestDof, 512);
    }
#undef UAAMG_CG_LAUNCH
    return hipGetLastError();
}

hipError_t finalize(int op, int n, const double* partials, PcgScalars* sc, HostMailbox* mb, hipStream_t s) {
    finalize_kernel<<<1, kFinThreads, 0, s>>>(op, n, partials, sc, mb);
    return hipGetLastError();
}

}  // namespace

int reduction_blocks(const LevelView& L) { return int(grid_for(L.nleaf)) > 0 ? int(grid_for(L.nleaf)) : 1; }

int coarsest_cg_threads() { return cg_threads(); }

hipError_t launch_coarsest_cg(const CoarsestView& C, const float* b, float* x, hipStream_t s) {
    if (C.ndof <= 0) return hipSuccess;
    if (C.ndof > kMaxCoarsestDof) return hipErrorInvalidValue;
    return launch_cg_variant<float>(false, nullptr, C, 1, C.ndof, b, x, s);
}

hipError_t launch_coarsest_cg(const CoarsestView& C, const bf16* b, bf16* x, hipStream_t s) {
    if (C.ndof <= 0) return hipSuccess;
    if (C.ndof > kMaxCoarsestDof) return hipErrorInvalidValue;
    return launch_cg_variant<bf16>(false, nullptr, C, 1, C.ndof, b, x, s);
}

hipError_t launch_residual64_absmax(const LevelView& L, const float* x, const float* b, float* r, PcgScalars* sc,
                                    double* partials, HostMailbox* mb, hipStream_t s) {
    const int g = reduction_blocks(L);
    residual64_kernel<<<unsigned(g), kBlockThreads, 0, s>>>(L, x, b, r, partials);
    if (hipError_t e = hipGetLastError()) return e;
    return finalize(kFinNu, g, partials, sc, mb, s);
}

hipError_t launch_replace64_absmax(const LevelView& L, const float* x, const float* p, const float* b, float* r,
                                  PcgScalars* sc, double* partials, HostMailbox* mb, hipStream_t s) {
    const int g = reduction_blocks(L);
    replace64_kernel<<<unsigned(g), kBlockThreads, 0, s>>>(L, x, p, b, r, sc, partials);
    if (hipError_t e = hipGetLastError()) return e;
    return finalize(kFinNu, g, partials, sc, mb, s);
}

hipError_t launch_dot(const LevelView& L, const float* a, const float* b, PcgScalars* sc, double* partials,
                      int target, hipStream_t s) {
    const int g = reduction_blocks(L);
    dot_kernel<0><<<unsigned(g), kBlockThreads, 0, s>>>(L, a, b, partials);
    if (hipError_t e = hipGetLastError()) return e;
    const int op = target == kDotRho ? kFinRho : (target == kDotSigma ? kFinSigma : kFinRhoNew);
    return finalize(op, g, partials, sc, nullptr, s);
}

hipError_t launch_laplacian_dot(const LevelView& L, const float* p, float* z, PcgScalars* sc, double* partials,
                                hipStream_t s) {
    const int g = reduction_blocks(L);
    laplacian_dot_kernel<<<unsigned(g), kBlockThreads, 0, s>>>(L, p, z, partials);
    if (hipError_t e = hipGetLastError()) return e;
    return finalize(kFinSigma, g, partials, sc, nullptr, s);
}

hipError_t launch_axpy_absmax(const LevelView& L, const float* z, float* r, PcgScalars* sc, double* partials,
                              HostMailbox* mb, hipStream_t s) {
    const int g = reduction_blocks(L);
    absmax_kernel<true><<<unsigned(g), kBlockThreads, 0, s>>>(L, z, r, sc, partials);
    if (hipError_t e = hipGetLastError()) return e;
    return finalize(kFinNu, g, partials, sc, mb, s);
}

hipError_t launch_absmax(const LevelView& L, const float* r, PcgScalars* sc, double* partials, HostMailbox* mb,
                         hipStream_t s) {
    const int g = reduction_blocks(L);
    absmax_kernel<false><<<unsigned(g), kBlockThreads, 0, s>>>(L, nullptr, const_cast<float*>(r), sc, partials);
    if (hipError_t e = hipGetLastError()) return e;
    return finalize(kFinNu, g, partials, sc, mb, s);
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

hipError_t launch_jacobi_dot(const LevelView& L, const float* x, const float* b, float* y, float w, PcgScalars* sc,
                             double* partials, int target, hipStream_t s) {
    const int g = reduction_blocks(L);
    jacobi_dot_kernel<float><<<unsigned(g), kBlockThreads, 0, s>>>(L, x, b, y, w, partials);
    if (hipError_t e = hipGetLastError()) return e;
    return finalize(target == kDotRho ? kFinRho : kFinRhoNew, g, partials, sc, nullptr, s);
}

hipError_t launch_jacobi_dot(const LevelView& L, const bf16* x, const float* b, float* y, float w, PcgScalars* sc,
                             double* partials, int target, hipStream_t s) {
    const int g = reduction_blocks(L);
    jacobi_dot_kernel<bf16><<<unsigned(g), kBlockThreads, 0, s>>>(L, x, b, y, w, partials);
    if (hipError_t e = hipGetLastError()) return e;
    return finalize(target == kDotRho ? kFinRho : kFinRhoNew, g, partials, sc, nullptr, s);
}

hipError_t launch_update_laplacian_dot(const LevelView& L, const float* z, const float* p, float* p_new, float* x,
                                       float* q, PcgScalars* sc, double* partials, hipStream_t s) {
    const int g = reduction_blocks(L);
    update_laplacian_dot_kernel<<<unsigned(g), kBlockThreads, 0, s>>>(L, z, p, p_new, x, q, sc, partials);
    if (hipError_t e = hipGetLastError()) return e;
    return finalize(kFinSigma, g, partials, sc, nullptr, s);
}

// ---- batched solve launchers (batch_kernels.h)
hipError_t launch_absmax_partials(const LevelView& L, const float* r, double* partials, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    absmax_kernel<false><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, nullptr, const_cast<float*>(r), nullptr,
                                                                     partials);
    return hipGetLastError();
}

hipError_t launch_residual64_partials(const LevelView& L, const float* x, const float* b, float* r, double* partials,
                                      hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    residual64_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, x, b, r, partials);
    return hipGetLastError();
}

hipError_t launch_dot_partials(const LevelView& L, const float* a, const float* b, double* partials, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    dot_kernel<0><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, a, b, partials);
    return hipGetLastError();
}

hipError_t launch_laplacian_dot_partials(const LevelView& L, const float* p, float* q, double* partials,
                                         hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    laplacian_dot_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, p, q, partials);
    return hipGetLastError();
}

hipError_t launch_jacobi_dot_partials(const LevelView& L, const float* x, const float* b, float* y, float w,
                                      double* partials, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    jacobi_dot_kernel<float><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, x, b, y, w, partials);
    return hipGetLastError();
}

hipError_t launch_jacobi_dot_partials(const LevelView& L, const bf16* x, const float* b, float* y, float w,
                                      double* partials, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    jacobi_dot_kernel<bf16><<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, x, b, y, w, partials);
    return hipGetLastError();
}

hipError_t launch_bupdate_laplacian_dot(const LevelView& L, const int32_t* leafprob, const BatchPcg* st,
                                        const float* z, const float* p, float* p_new, float* x, float* q,
                                        double* partials, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    bupdate_laplacian_dot_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, leafprob, st, z, p, p_new, x, q,
                                                                            partials);
    return hipGetLastError();
}

hipError_t launch_baxpy_absmax(const LevelView& L, const int32_t* leafprob, const BatchPcg* st, const float* q,
                               const float* x, const float* p, const float* b, float* r, double* partials,
                               hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    baxpy_absmax_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, leafprob, st, q, x, p, b, r, partials);
    return hipGetLastError();
}

hipError_t launch_bupdate_x(const LevelView& L, const int32_t* leafprob, BatchPcg* st, const float* p, float* x,
                            hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    bupdate_x_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, leafprob, st, p, x);
    if (hipError_t e = hipGetLastError()) return e;
    return hipSuccess;
}

hipError_t launch_bupdate_xp(const LevelView& L, const int32_t* leafprob, const BatchPcg* st, const float* z,
                             float* x, float* p, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    bupdate_xp_kernel<<<grid_for(L.nleaf), kBlockThreads, 0, s>>>(L, leafprob, st, z, x, p);
    return hipGetLastError();
}

hipError_t launch_bfinalize(int op, int nprob, const int32_t* blk_begin, const double* partials, BatchPcg* st,
                            hipStream_t s) {
    if (nprob <= 0) return hipSuccess;
    if (op < 0) {   // commit: converging -> success
        bcommit_kernel<<<1, kFinThreads, 0, s>>>(nprob, st);
        return hipGetLastError();
    }
    bfinalize_kernel<<<unsigned(nprob), kFinThreads, 0, s>>>(op, blk_begin, partials, st);
    return hipGetLastError();
}

hipError_t launch_bstatus(int nprob, const BatchPcg* st, unsigned int* seq, HostMailbox* mb, hipStream_t s) {
    bstatus_kernel<<<1, kFinThreads, 0, s>>>(nprob, st, seq, mb);
    return hipGetLastError();
}

hipError_t launch_bcoarsest(const CoarsestView* views, int n, int max_ndof, const float* b, float* x, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (max_ndof > kMaxCoarsestDof) return hipErrorInvalidValue;
    return launch_cg_variant<float>(true, views, CoarsestView{}, unsigned(n), max_ndof, b, x, s);
}

hipError_t launch_bcoarsest(const CoarsestView* views, int n, int max_ndof, const bf16* b, bf16* x, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (max_ndof > kMaxCoarsestDof) return hipErrorInvalidValue;
    return launch_cg_variant<bf16>(true, views, CoarsestView{}, unsigned(n), max_ndof, b, x, s);
}

}  // namespace uaamg
