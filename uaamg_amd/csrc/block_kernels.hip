// One W-cycle level visit in two launches (coarse levels).
//
// On the small levels of the hierarchy every smoothing sweep is its own
// latency-bound launch. A block-visit kernel gives each leaf one workgroup
// that stages the leaf plus a halo of H voxels (from up to 26 neighbour
// leaves) in LDS and runs consecutive sweeps on a shrinking region: sweep k
// is valid on [k, S-k)^3 of the S = 8 + 2H region, so after H-1 sweeps plus
// the residual (or H sweeps) the leaf's own 8^3 values are exact. The halo
// voxels are recomputed redundantly by each neighbouring workgroup; every
// value is produced by the same arithmetic as the per-sweep kernels
// (stencil order ((x+ + x-) + (y+ + y-)) + (z+ + z-), fma(x, diag, s),
// fma(w, (b - Ax) * D^-1, x), non-DOF voxels 0), and with BF16 storage each
// sweep's output is rounded where the per-sweep kernels store it, so results
// are bit-identical to theirs in either precision.
//
//   kPreZero  (H = 3): x1 = w0 D^-1 b, two sweeps, residual, restriction
//             (first visit, muCyclePreconditioner from a zero guess)
//   kPreAgain (H = 4): three sweeps from the level's x, residual, restriction
//   kPost     (H = 3): x + P x_coarse, three sweeps
//
// On the level above the coarsest, the post-visit can take the coarsest solve
// in with it (block_post_cg_kernel): every workgroup runs the coarsest CG on
// the whole coarsest system (the same routine and thread count as the
// standalone solve, so every workgroup gets the same values the standalone
// kernel stores), keeps the solution in LDS for its staging, and workgroup 0
// also stores it. The coarse correction then never crosses HBM and the
// W-cycle bottom loses a launch.
#include <type_traits>

#include "coarsest_device.h"
#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

enum BlockMode { kPreZero = 0, kPreAgain = 1, kPost = 2 };
constexpr int kBlkThreads = 512;

// LDS per region voxel: b, x, diag, invdiag, three face coefficients and the
// mask byte (29 bytes; sweeps update x in place through registers), so two
// H = 3 workgroups fit one CU's 160 KiB.
template <int H>
struct Region {
    static constexpr int S = 8 + 2 * H;
    static constexpr int N = S * S * S;
    static constexpr size_t lds_bytes = size_t(N) * (7 * sizeof(float) + 1);
};

template <int H>
__device__ __forceinline__ int ridx(int x, int y, int z) {
    return (x * Region<H>::S + y) * Region<H>::S + z;
}

// value as the vector storage type holds it (the per-sweep kernels store
// every sweep's output, so the block form rounds where they would)
template <class V>
__device__ __forceinline__ float stored(float v) {
    if constexpr (std::is_same_v<V, float>) return v;
    else return float(V(v));
}

// One Jacobi sweep x <- J_w(x) on [LO, S - LO)^3 of the region, each output
// rounded to the storage type V: every thread computes its voxels into
// registers, then, after a barrier, stores them (the caller's barrier
// publishes the stores).
template <int H, int LO, class V>
__device__ __forceinline__ void sweep(float* x, const float* b, const float* dg, const float* inv, const float* cx,
                                      const float* cy, const float* cz, const uint8_t* m, float w) {
    constexpr int S = Region<H>::S;
    constexpr int e = S - 2 * LO;
    constexpr int cnt = e * e * e;
    constexpr int kPer = (cnt + kBlkThreads - 1) / kBlkThreads;
    float out[kPer];
    int at[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int q = threadIdx.x + k * kBlkThreads;
        at[k] = -1;
        out[k] = 0.f;
        if (q < cnt) {
            const int vx = LO + q / (e * e), vy = LO + (q / e) % e, vz = LO + q % e;
            const int i = ridx<H>(vx, vy, vz);
            at[k] = i;
            float r = 0.f;
            if (m[i]) {
                const int ix = S * S, iy = S;
                const float f1 = x[i + ix] * cx[i + ix];
                const float f2 = x[i - ix] * cx[i];
                const float f3 = x[i + iy] * cy[i + iy];
                const float f4 = x[i - iy] * cy[i];
                const float f5 = x[i + 1] * cz[i + 1];
                const float f6 = x[i - 1] * cz[i];
                const float s = ((f1 + f2) + (f3 + f4)) + (f5 + f6);
                const float ax = fmaf(x[i], dg[i], s);
                float t = b[i] - ax;
                t = t * inv[i];
                r = fmaf(w, t, x[i]);
            }
            out[k] = stored<V>(r);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (at[k] >= 0) x[at[k]] = out[k];
}

// The visit of one leaf. kCg (kPost only): the coarse correction comes from
// cgx, the coarsest solution in LDS indexed by coarsest DOF (dof_of maps a
// coarsest brick position to its DOF, -1 elsewhere), instead of Xc.
template <int H, int MODE, class V, bool kCg>
__device__ __forceinline__ void visit_body(const LevelView& L, const V* Xin, const V* Xc, const V* B, V* Xout, V* Cv,
                                           float w0, float w1, float w2, const int32_t* dof_of, const float* cgx) {
    constexpr int S = Region<H>::S, N = Region<H>::N;
    extern __shared__ float lds[];
    float* b = lds;
    float* xa = b + N;
    float* dg = xa + N;
    float* inv = dg + N;
    float* cx = inv + N;
    float* cy = cx + N;
    float* cz = cy + N;
    uint8_t* m = reinterpret_cast<uint8_t*>(cz + N);
    __shared__ int nl[27], sd[27], sx[27], sy[27], sz[27], par[27];

    const int leaf = logical_block();
    const int tid = threadIdx.x;
    if (tid < 27) {
        const int4* r = reinterpret_cast<const int4*>(L.nb27 + (size_t(leaf) * 27 + tid) * kNb27Stride);
        const int4 a = r[0], b = r[1];
        nl[tid] = a.x;
        sd[tid] = a.y;
        sx[tid] = a.z;
        sy[tid] = a.w;
        sz[tid] = b.x;
        par[tid] = MODE == kPost ? b.y : -1;
    }
    __syncthreads();

    // stage the region: values of absent leaves are 0 (vectors, mask) or the
    // default stencil constants (coefficients), as the per-sweep kernels read them
    // unrolled so that every thread's loads are in flight together (the
    // region's voxels are independent; one memory latency instead of one per
    // pass of the loop)
    constexpr int kPasses = (N + kBlkThreads - 1) / kBlkThreads;
#pragma unroll
    for (int pass = 0; pass < kPasses; ++pass) {
        const int q = tid + pass * kBlkThreads;
        if (q >= N) break;
        const int gx = q / (S * S) - H, gy = (q / S) % S - H, gz = q % S - H;
        const int ax = gx < 0 ? 0 : (gx > 7 ? 2 : 1), ay = gy < 0 ? 0 : (gy > 7 ? 2 : 1);
        const int az = gz < 0 ? 0 : (gz > 7 ? 2 : 1);
        const int k = ax * 9 + ay * 3 + az;
        const int n = nl[k];
        const int lx = gx & 7, ly = gy & 7, lz = gz & 7;
        const int off = (lx << 6) | (ly << 3) | lz;
        // every load is unconditional (absent leaf or slot: leaf 0 / slot 0,
        // both exist, and the value is replaced below), so the loads of all
        // passes are in flight together
        const size_t nv = size_t(n >= 0 ? n : 0);
        const int s0 = sd[k], s1 = sx[k], s2 = sy[k], s3 = sz[k];
        const float tdg = L.diag[size_t(s0 >= 0 ? s0 : 0) * 512 + off];
        const float tinv = L.invdiag[size_t(s0 >= 0 ? s0 : 0) * 512 + off];
        const float tcx = L.cx[size_t(s1 >= 0 ? s1 : 0) * 512 + off];
        const float tcy = L.cy[size_t(s2 >= 0 ? s2 : 0) * 512 + off];
        const float tcz = L.cz[size_t(s3 >= 0 ? s3 : 0) * 512 + off];
        uint64_t tm = L.mask[nv * 8 + lx];
        float tb = float(B[nv * 512 + off]);
        float tx = 0.f, tc = 0.f;
        int pp = -1, di = -1;
        if (MODE != kPreZero) tx = float(Xin[nv * 512 + off]);
        if (MODE == kPost) {
            pp = par[k];
            const int p = pp >= 0 ? pp : 0, cl = p >> 3, oct = p & 7;
            const int crow = ((lx >> 1) + 4 * ((oct >> 2) & 1)) * 8 + ((ly >> 1) + 4 * ((oct >> 1) & 1));
            const size_t ci = size_t(cl) * 512 + crow * 8 + 4 * (oct & 1) + (lz >> 1);
            if constexpr (kCg) di = dof_of[ci];
            else tc = float(Xc[ci]);
        }
        // selects, no branches: the empty asm statements (no instruction)
        // keep the compiler from sinking these loads into a branch on n
        asm volatile("" : "+v"(tm), "+v"(tb), "+v"(tx), "+v"(tc), "+v"(di));
        // a fine DOF's parent voxel is a coarsest DOF, the only case where tc is used
        if constexpr (kCg) tc = cgx[di >= 0 ? di : kMaxCoarsestDof];
        const bool has = n >= 0;
        const float vdg = s0 >= 0 ? tdg : L.dflt_diag;
        const float vinv = s0 >= 0 ? tinv : L.dflt_invdiag;
        const uint8_t vm = has ? uint8_t((tm >> ((ly << 3) | lz)) & 1u) : uint8_t(0);
        const float vb = has ? tb : 0.f;
        float vx;
        if (MODE == kPreZero) {
            const float t = w0 * (vinv * vb);
            vx = has ? t : 0.f;
        } else {
            vx = has ? tx : 0.f;
            if (MODE == kPost) {
                const float t = vx + tc;
                vx = (vm && pp >= 0) ? t : vx;
            }
        }
        b[q] = vb;
        xa[q] = vx;
        m[q] = vm;
        dg[q] = vdg;
        inv[q] = vinv;
        cx[q] = s1 >= 0 ? tcx : L.dflt_off;
        cy[q] = s2 >= 0 ? tcy : L.dflt_off;
        cz[q] = s3 >= 0 ? tcz : L.dflt_off;
    }
    __syncthreads();

    if (MODE == kPreZero) {
        sweep<H, 1, V>(xa, b, dg, inv, cx, cy, cz, m, w1);
        __syncthreads();
        sweep<H, 2, V>(xa, b, dg, inv, cx, cy, cz, m, w2);
        __syncthreads();
    } else {
        sweep<H, 1, V>(xa, b, dg, inv, cx, cy, cz, m, w0);
        __syncthreads();
        sweep<H, 2, V>(xa, b, dg, inv, cx, cy, cz, m, w1);
        __syncthreads();
        sweep<H, 3, V>(xa, b, dg, inv, cx, cy, cz, m, w2);
        __syncthreads();
    }
    const float* fin = xa;   // the final sweep

    // the leaf's own values
    for (int q = tid; q < 512; q += kBlkThreads) {
        const int i = ridx<H>(H + (q >> 6), H + ((q >> 3) & 7), H + (q & 7));
        Xout[size_t(leaf) * 512 + q] = V(fin[i]);
    }
    if (MODE == kPost) return;

    // residual on the own brick (into registers; then the first 512 entries
    // of the diagonal array, no longer read, hold it), then restriction into
    // the parent's octant
    static_assert(kBlkThreads == 512, "one own-brick voxel per thread");
    float rv = 0.f;
    {
        const int q = tid;
        const int i = ridx<H>(H + (q >> 6), H + ((q >> 3) & 7), H + (q & 7));
        if (m[i]) {
            const int ix = S * S, iy = S;
            const float* x = fin;
            const float f1 = x[i + ix] * cx[i + ix];
            const float f2 = x[i - ix] * cx[i];
            const float f3 = x[i + iy] * cy[i + iy];
            const float f4 = x[i - iy] * cy[i];
            const float f5 = x[i + 1] * cz[i + 1];
            const float f6 = x[i - 1] * cz[i];
            const float s = ((f1 + f2) + (f3 + f4)) + (f5 + f6);
            rv = b[i] - fmaf(x[i], dg[i], s);
        }
    }
    __syncthreads();
    float* rs = dg;
    rs[tid] = rv;
    __syncthreads();
    if (tid < 64) {
        const int ccx = tid >> 4, ccy = (tid >> 2) & 3, ccz = tid & 3;
        float sum = 0.f;
        for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj)
                for (int kk = 0; kk < 2; ++kk) sum += rs[((2 * ccx + ii) * 8 + (2 * ccy + jj)) * 8 + 2 * ccz + kk];
        const int p = L.parent[leaf];
        if (p < 0) return;
        const int cl = p >> 3, oct = p & 7;
        const int off = ((ccx + 4 * ((oct >> 2) & 1)) * 8 + (ccy + 4 * ((oct >> 1) & 1))) * 8 + ccz + 4 * (oct & 1);
        Cv[size_t(cl) * 512 + off] = V(sum * 0.125f);
    }
}

template <int H, int MODE, class V>
__global__ __launch_bounds__(kBlkThreads) void block_visit_kernel(LevelView L, const V* Xin, const V* Xc, const V* B,
                                                                  V* Xout, V* Cv, float w0, float w1, float w2) {
    visit_body<H, MODE, V, false>(L, Xin, Xc, B, Xout, Cv, w0, w1, w2, nullptr, nullptr);
}

// The coarsest solve (right-hand side Bc, solution stored into Xc by
// workgroup 0) followed by the kPost visit of the level above it. The CG's
// last barrier has passed when it writes its solution into cg.sp, and the
// visit's first barrier publishes it.
template <class V, int kMaxDof>
__global__ __launch_bounds__(kBlkThreads) void block_post_cg_kernel(LevelView L, CoarsestView C,
                                                                    const int32_t* dof_of, const V* Bc, V* Xc,
                                                                    const V* Xin, const V* B, V* Xout, float w0,
                                                                    float w1, float w2) {
    __shared__ CgShared<kBlkThreads> cg;
    coarsest_cg_block<kBlkThreads, V, kMaxDof, true>(C, Bc, blockIdx.x == 0 ? Xc : nullptr, cg);
    visit_body<3, kPost, V, true>(L, Xin, nullptr, B, Xout, nullptr, w0, w1, w2, dof_of, cg.sp);
}

template <int H, int MODE, class V>
hipError_t launch_mode(const LevelView& L, const V* xin, const V* xc, const V* b, V* xout, V* cv, float w0, float w1,
                       float w2, hipStream_t s) {
    block_visit_kernel<H, MODE, V><<<unsigned(L.nleaf), kBlkThreads, Region<H>::lds_bytes, s>>>(L, xin, xc, b, xout,
                                                                                                cv, w0, w1, w2);
    return hipGetLastError();
}

template <class V>
hipError_t set_lds_limits() {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&block_visit_kernel<3, kPreZero, V>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(Region<3>::lds_bytes));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&block_visit_kernel<4, kPreAgain, V>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(Region<4>::lds_bytes));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&block_visit_kernel<3, kPost, V>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(Region<3>::lds_bytes));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&block_post_cg_kernel<V, kCgSmallDof>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(Region<3>::lds_bytes));
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&block_post_cg_kernel<V, kMaxCoarsestDof>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(Region<3>::lds_bytes));
    return e;
}

template <class V>
hipError_t block_pre_t(const LevelView& L, bool from_zero, const V* x, const V* b, V* x_out, V* coarse_b, float w0,
                       float w1, float w2, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    if (!L.nb27) return hipErrorInvalidValue;
    if (from_zero) return launch_mode<3, kPreZero, V>(L, nullptr, nullptr, b, x_out, coarse_b, w0, w1, w2, s);
    return launch_mode<4, kPreAgain, V>(L, x, nullptr, b, x_out, coarse_b, w0, w1, w2, s);
}

template <class V>
hipError_t block_post_t(const LevelView& L, const V* x, const V* xc, const V* b, V* x_out, float w0, float w1, float w2,
                        hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;
    if (!L.nb27) return hipErrorInvalidValue;
    return launch_mode<3, kPost, V>(L, x, xc, b, x_out, nullptr, w0, w1, w2, s);
}

template <class V>
hipError_t block_post_cg_t(const LevelView& L, const CoarsestView& C, const int32_t* dof_of, const V* bc, V* xc,
                           const V* x, const V* b, V* x_out, float w0, float w1, float w2, hipStream_t s) {
    if (L.nleaf <= 0) return hipSuccess;   // as block_post_t: nothing reads the coarse correction
    if (!L.nb27 || !dof_of || C.ndof <= 0 || C.ndof > kMaxCoarsestDof) return hipErrorInvalidValue;
    const unsigned g = unsigned(L.nleaf);
    const size_t lds = Region<3>::lds_bytes;
    if (C.ndof <= kCgSmallDof)
        block_post_cg_kernel<V, kCgSmallDof><<<g, kBlkThreads, lds, s>>>(L, C, dof_of, bc, xc, x, b, x_out, w0, w1, w2);
    else
        block_post_cg_kernel<V, kMaxCoarsestDof>
            <<<g, kBlkThreads, lds, s>>>(L, C, dof_of, bc, xc, x, b, x_out, w0, w1, w2);
    return hipGetLastError();
}

}  // namespace

hipError_t block_kernels_init() {
    hipError_t e = set_lds_limits<float>();
    if (e == hipSuccess) e = set_lds_limits<bf16>();
    return e;
}

hipError_t launch_block_pre(const LevelView& L, bool from_zero, const float* x, const float* b, float* x_out,
                            float* coarse_b, float w0, float w1, float w2, hipStream_t s) {
    return block_pre_t<float>(L, from_zero, x, b, x_out, coarse_b, w0, w1, w2, s);
}

hipError_t launch_block_post(const LevelView& L, const float* x, const float* xc, const float* b, float* x_out,
                             float w0, float w1, float w2, hipStream_t s) {
    return block_post_t<float>(L, x, xc, b, x_out, w0, w1, w2, s);
}

hipError_t launch_block_pre(const LevelView& L, bool from_zero, const bf16* x, const bf16* b, bf16* x_out,
                            bf16* coarse_b, float w0, float w1, float w2, hipStream_t s) {
    return block_pre_t<bf16>(L, from_zero, x, b, x_out, coarse_b, w0, w1, w2, s);
}

hipError_t launch_block_post(const LevelView& L, const bf16* x, const bf16* xc, const bf16* b, bf16* x_out,
                             float w0, float w1, float w2, hipStream_t s) {
    return block_post_t<bf16>(L, x, xc, b, x_out, w0, w1, w2, s);
}

hipError_t launch_block_post_cg(const LevelView& L, const CoarsestView& C, const int32_t* dof_of, const float* bc,
                                float* xc, const float* x, const float* b, float* x_out, float w0, float w1, float w2,
                                hipStream_t s) {
    return block_post_cg_t<float>(L, C, dof_of, bc, xc, x, b, x_out, w0, w1, w2, s);
}

hipError_t launch_block_post_cg(const LevelView& L, const CoarsestView& C, const int32_t* dof_of, const bf16* bc,
                                bf16* xc, const bf16* x, const bf16* b, bf16* x_out, float w0, float w1, float w2,
                                hipStream_t s) {
    return block_post_cg_t<bf16>(L, C, dof_of, bc, xc, x, b, x_out, w0, w1, w2, s);
}

}  // namespace uaamg
