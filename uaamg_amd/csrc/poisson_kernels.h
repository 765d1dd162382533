// Device kernels of the UAAMG-PCG Poisson solve on 8^3 leaf bricks (gfx950).
//
// Layout in HBM, per level: a level vector is [nleaf][512] values in leaf
// (tree) order with voxel offset 64x+8y+z; entries that are not DOFs hold 0.
// Every kernel maps one wavefront to one leaf and one lane to one z-row
// (x,y) = (lane>>3, lane&7) of 8 contiguous values (32 bytes in FP32), so a
// wavefront reads and writes its brick as one contiguous 2 KiB block.
// Coefficient bricks exist only for leaves the reference keeps after
// trimming; slot[leaf][k] = -1 selects the default stencil constant.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace uaamg {

// storage type of the W-cycle vectors in the mixed-precision form (arithmetic stays FP32)
using bf16 = __bf16;

// Per-leaf record of 32 int32 (128 bytes, scalar loads issued together at
// the top of a kernel): the coefficient-brick slots of the leaf, the slots
// the +x/+y/+z faces read from the neighbour leaf (-1: neighbour absent or
// its brick trimmed, both read the default constant), the six face
// neighbours, the leaf's parent (coarse leaf * 8 + octant, -1 on the coarsest
// level) and each face neighbour's diagonal slot and parent. With them every
// data load of a stencil kernel depends on one memory round trip only.
enum {
    kSlotDiag = 0,
    kSlotX = 1,
    kSlotY = 2,
    kSlotZ = 3,
    kSlotSpai0 = 4,
    kSlotXp = 5,          // kSlotX of the +x neighbour
    kSlotYp = 6,          // kSlotY of the +y neighbour
    kSlotZp = 7,          // kSlotZ of the +z neighbour
    kSlotNb = 8,          // 8..13: face neighbours -x,+x,-y,+y,-z,+z (as nbr)
    kSlotParent = 14,
    kSlotNbDiag = 16,     // 16..21: kSlotDiag of each face neighbour
    kSlotNbParent = 22,   // 22..27: parent of each face neighbour
    kSlotStride = 32
};

// Fills the neighbour and parent fields of a record table whose own-slot
// fields (kSlotDiag .. kSlotSpai0) are set; nbr is [n][6], parent [n] or null
// (not known yet: -1).
inline void complete_leaf_records(int32_t* rec, const int32_t* nbr, const int32_t* parent, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        int32_t* r = rec + i * kSlotStride;
        const int32_t* nb = nbr + i * 6;
        r[kSlotXp] = nb[1] >= 0 ? rec[size_t(nb[1]) * kSlotStride + kSlotX] : -1;
        r[kSlotYp] = nb[3] >= 0 ? rec[size_t(nb[3]) * kSlotStride + kSlotY] : -1;
        r[kSlotZp] = nb[5] >= 0 ? rec[size_t(nb[5]) * kSlotStride + kSlotZ] : -1;
        r[kSlotParent] = parent ? parent[i] : -1;
        r[15] = -1;
        for (int f = 0; f < 6; ++f) {
            r[kSlotNb + f] = nb[f];
            r[kSlotNbDiag + f] = nb[f] >= 0 ? rec[size_t(nb[f]) * kSlotStride + kSlotDiag] : -1;
            r[kSlotNbParent + f] = (parent && nb[f] >= 0) ? parent[nb[f]] : -1;
        }
        for (int k = 28; k < kSlotStride; ++k) r[k] = -1;
    }
}

// Neighbourhood records of the block-visit kernels: per leaf and offset
// (ax,ay,az) in {-1,0,1}^3 (k = (ax+1)*9 + (ay+1)*3 + az+1) the neighbour leaf
// (-1 none), its diagonal / x / y / z slots and its parent (-1 when the leaf
// is absent), so a workgroup's staging depends on one memory round trip.
constexpr int kNb27Stride = 8;

// nb27: [n][27] neighbour leaves; rec: the level's leaf records; parent [n] or null.
inline void build_nb27_records(int32_t* out, const int32_t* nb27, const int32_t* rec, const int32_t* parent,
                               size_t n) {
    for (size_t i = 0; i < n * 27; ++i) {
        const int32_t m = nb27[i];
        int32_t* r = out + i * kNb27Stride;
        r[0] = m;
        for (int k = 0; k < 4; ++k) r[1 + k] = m >= 0 ? rec[size_t(m) * kSlotStride + size_t(kSlotDiag + k)] : -1;
        r[5] = (m >= 0 && parent) ? parent[m] : -1;
        r[6] = r[7] = -1;
    }
}

// smoothing / apply modes of LaplacianApplySIMD (simd_vdb_poisson.cpp:683-690)
enum ApplyMode : int {
    kModeLaplacian = 0,
    kModeResidual = 1,
    kModeJacobi = 2,
    kModeSpai0 = 3,
    kModeRedGS = 4,
    kModeBlackGS = 5,
    kModeJacobiBf16 = 6,   // measurement only: the weighted Jacobi sweep of the BF16-stored W-cycle
    kModeBlockPost = 7,    // measurement only: a coarse level's post-visit block kernel (prolongation + 3 sweeps)
    kModeCoarsest = 8,     // measurement only: the coarsest-level CG (level = the coarsest)
};

struct LevelView {
    int32_t nleaf;
    int32_t level;            // 0 selects the level-0 kernel instantiations (separate profile names)
    float term;
    const uint64_t* mask;     // [nleaf][8]
    const int32_t* nbr;       // [nleaf][6]  -x,+x,-y,+y,-z,+z
    const int32_t* slot;      // [nleaf][kSlotStride] leaf records
    const int64_t* dofbase;   // [nleaf]
    const int32_t* parent;    // [nleaf] coarse leaf * 8 + octant (fine side of a transfer)
    const int32_t* child;     // [nleaf][8] fine leaf per octant (coarse side of a transfer)
    const float* diag;        // [nslot][512]
    const float* invdiag;     // [nslot][512], same slots as diag
    const float* cx;
    const float* cy;
    const float* cz;
    const float* spai0;
    float dflt_diag, dflt_off, dflt_invdiag, dflt_spai0;
    const int32_t* nb27;      // [nleaf][27][kNb27Stride] neighbourhood records (block-visit kernels; null on level 0)
};

// Scalars of the outer PCG, kept in device memory so a whole iteration runs
// without host round trips except for the convergence test.
struct PcgScalars {
    double rho;        // <z, r> of the current direction
    double sigma;      // <p, A p>
    double rho_new;
    double dot_tmp;
    float alpha;
    float beta;
    float nu;          // ||r||_inf (NaN if any entry is not finite)
    unsigned int seq;  // number of completed ||r||_inf reductions
    unsigned int counter[8];   // completion counters of the reduction kernels
};

// Host-mapped pinned mailbox: each ||r||_inf reduction publishes
// (seq << 32 | bits of nu) here in one 64-bit store, so the host polls one
// word without a stream synchronisation and cannot see a torn pair.
struct alignas(64) HostMailbox {
    volatile uint64_t word;
    uint64_t pad[7];
};

// Compact coarsest-level system: per DOF the brick position, 6 neighbour DOFs
// and the 7 stencil values in the operator's summation order.
struct CoarsestView {
    int32_t ndof;
    const int32_t* pos;       // [ndof] leaf*512 + offset
    const int32_t* nb;        // [ndof][6]: +x,-x,+y,-y,+z,-z DOF index or -1
    const float* coef;        // [ndof][8]: diag, c+x, c-x, c+y, c-y, c+z, c-z, invdiag (Eigen's 1/diag or 1)
    int32_t max_iters;        // 10 (constructCoarsestLevelExactSolver)
    float tol;                // Eigen default: float epsilon
};

// Bottom of the W-cycle in one launch: nvisits visits of level L-2 (the first
// from a zero guess when first_skip) with the coarsest solve inside.
struct TailArgs {
    LevelView L;          // level L-2
    const float* B;       // its right-hand side
    float* X;             // its solution (final sweep output)
    float* T;             // smoothing buffers
    float* U;
    float* Bc;            // coarsest right-hand side and solution
    float* Xc;
    CoarsestView C;
    float wpre[3];        // Jacobi weights of the pre sweeps; post sweeps use wpost
    float wpost[3];
    int nvisits;
    int first_skip;
};

constexpr int kWavesPerBlock = 4;
constexpr int kBlockThreads = 64 * kWavesPerBlock;
constexpr int kMaxCoarsestDof = 4096;
// coarsest systems up to this size keep four row slots per thread of a
// 512-thread solve (fewer registers; about 8% faster at ~1700 DOF)
constexpr int kCgSmallDof = 2048;

// ---- launchers (stream-ordered, no host synchronisation) ----
// y = op(x, b) on every DOF leaf of a level; w = Jacobi / SOR weight.
hipError_t launch_apply(const LevelView& L, int mode, const float* x, const float* b, float* y, float w,
                        hipStream_t s);
// first sweep from a zero guess: 0 = Jacobi (y = w*D^-1 b), 1 = SPAI0, 2 = red SOR part of RBGS
hipError_t launch_first_sweep(const LevelView& L, int kind, const float* b, float* y, float w, hipStream_t s);
// coarse = 1/8 * sum of on fine children (restriction, simd_vdb_poisson.cpp:1624-1684)
hipError_t launch_restrict(const LevelView& coarse, const float* fine_v, float* coarse_v, hipStream_t s);
// fine += alpha * coarse on on fine voxels (prolongation<true>, :1686-1754)
hipError_t launch_prolong_add(const LevelView& fine, const float* coarse_v, float* fine_v, float alpha,
                              hipStream_t s);
// coarsest 10-iteration Jacobi-PCG from zero, one workgroup
hipError_t launch_coarsest_cg(const CoarsestView& C, const float* b, float* x, hipStream_t s);
// the last two levels of a W-cycle visit pair, one workgroup
hipError_t launch_tail(const TailArgs& a, hipStream_t s);

// DOF-ordered vector <-> bricks
hipError_t launch_scatter_dof(const LevelView& L, const float* dofvec, float* bricks, hipStream_t s);
hipError_t launch_gather_dof(const LevelView& L, const float* bricks, float* dofvec, hipStream_t s);

// outer PCG vector kernels on level 0 (partials: >= reduction_blocks(L) doubles)
int reduction_blocks(const LevelView& L);
// which scalar a dot kernel finalises
enum DotTarget { kDotRho = 0, kDotRhoNew = 1, kDotSigma = 2 };   // kDotSigma also sets alpha = rho / sigma
hipError_t launch_dot(const LevelView& L, const float* a, const float* b, PcgScalars* sc, double* partials,
                      int target, hipStream_t s);
// z = A p and sigma = <p, z>; finalises alpha = rho / sigma
hipError_t launch_laplacian_dot(const LevelView& L, const float* p, float* z, PcgScalars* sc, double* partials,
                                hipStream_t s);
// r -= alpha z ; nu = ||r||_inf
hipError_t launch_axpy_absmax(const LevelView& L, const float* z, float* r, PcgScalars* sc, double* partials,
                              HostMailbox* mb, hipStream_t s);
// nu = ||r||_inf only
hipError_t launch_absmax(const LevelView& L, const float* r, PcgScalars* sc, double* partials, HostMailbox* mb,
                         hipStream_t s);
// x += alpha p ; p = z + beta p
hipError_t launch_update_xp(const LevelView& L, const float* z, float* x, float* p, const PcgScalars* sc,
                            hipStream_t s);
// ---- fused forms used by the W-cycle and the PCG loop (same arithmetic as the
// unfused kernels, fewer passes over HBM and fewer launches)
// first two Jacobi sweeps from a zero guess: y = J_w1(w0 D^-1 b)
hipError_t launch_jacobi_first2(const LevelView& L, const float* b, float* y, float w0, float w1, hipStream_t s);
// y = J_w(x + P x_coarse)
hipError_t launch_prolong_jacobi(const LevelView& L, const float* x, const float* xc, const float* b, float* y,
                                 float w, hipStream_t s);
// coarse_b = R (b - A x)
hipError_t launch_residual_restrict(const LevelView& fine, const float* x, const float* b, float* coarse_b,
                                    hipStream_t s);
// y = J_w(x) and <y, b> finalised into rho (kDotRho) or rho_new / beta (kDotRhoNew)
hipError_t launch_jacobi_dot(const LevelView& L, const float* x, const float* b, float* y, float w, PcgScalars* sc,
                             double* partials, int target, hipStream_t s);
// x += alpha p; p_new = z + beta p; q = A p_new; sigma; alpha = rho / sigma
hipError_t launch_update_laplacian_dot(const LevelView& L, const float* z, const float* p, float* p_new, float* x,
                                       float* q, PcgScalars* sc, double* partials, hipStream_t s);
// x += alpha p only (final step)
hipError_t launch_update_x(const LevelView& L, const float* p, float* x, const PcgScalars* sc, hipStream_t s);

// ---- block-visit kernels (one workgroup per leaf, halo staged in LDS; needs L.nb27)
hipError_t block_kernels_init();
// pre-smoothing (3 weighted Jacobi sweeps w0, w1, w2; from a zero guess when
// from_zero, else from x) into x_out, and coarse_b = R (b - A x_out)
hipError_t launch_block_pre(const LevelView& L, bool from_zero, const float* x, const float* b, float* x_out,
                            float* coarse_b, float w0, float w1, float w2, hipStream_t s);
// x_out = three sweeps (w0, w1, w2) from x + P x_coarse
hipError_t launch_block_post(const LevelView& L, const float* x, const float* xc, const float* b, float* x_out,
                             float w0, float w1, float w2, hipStream_t s);
// The coarsest solve xc = CG(bc) (as launch_coarsest_cg stores it) and then
// launch_block_post of the level above the coarsest with that xc, in one
// launch. dof_of: [coarsest leaves * 512], the coarsest DOF at each brick
// position or -1. Requires the coarsest CG at 512 threads.
hipError_t launch_block_post_cg(const LevelView& L, const CoarsestView& C, const int32_t* dof_of, const float* bc,
                                float* xc, const float* x, const float* b, float* x_out, float w0, float w1, float w2,
                                hipStream_t s);
hipError_t launch_block_post_cg(const LevelView& L, const CoarsestView& C, const int32_t* dof_of, const bf16* bc,
                                bf16* xc, const bf16* x, const bf16* b, bf16* x_out, float w0, float w1, float w2,
                                hipStream_t s);
// threads of the coarsest CG workgroup (512 unless UAAMG_CG_THREADS says otherwise)
int coarsest_cg_threads();

// ---- mixed-precision W-cycle: the same passes with the level vectors stored
// in BF16 and every operation computed in FP32. Level 0's right-hand side is
// the outer PCG's FP32 residual and its last sweep writes the FP32 z; every
// other vector of the cycle is BF16.
hipError_t launch_jacobi_first2(const LevelView& L, const float* b, bf16* y, float w0, float w1, hipStream_t s);
hipError_t launch_jacobi_first2(const LevelView& L, const bf16* b, bf16* y, float w0, float w1, hipStream_t s);
// y = J_w(x) (weighted Jacobi sweep)
hipError_t launch_jacobi(const LevelView& L, const bf16* x, const float* b, bf16* y, float w, hipStream_t s);
hipError_t launch_jacobi(const LevelView& L, const bf16* x, const float* b, float* y, float w, hipStream_t s);
hipError_t launch_jacobi(const LevelView& L, const float* x, const float* b, bf16* y, float w, hipStream_t s);
hipError_t launch_jacobi(const LevelView& L, const bf16* x, const bf16* b, bf16* y, float w, hipStream_t s);
hipError_t launch_residual_restrict(const LevelView& fine, const bf16* x, const float* b, bf16* coarse_b,
                                    hipStream_t s);
hipError_t launch_residual_restrict(const LevelView& fine, const bf16* x, const bf16* b, bf16* coarse_b,
                                    hipStream_t s);
hipError_t launch_prolong_jacobi(const LevelView& L, const bf16* x, const bf16* xc, const float* b, bf16* y, float w,
                                 hipStream_t s);
hipError_t launch_prolong_jacobi(const LevelView& L, const bf16* x, const bf16* xc, const bf16* b, bf16* y, float w,
                                 hipStream_t s);
hipError_t launch_jacobi_dot(const LevelView& L, const bf16* x, const float* b, float* y, float w, PcgScalars* sc,
                             double* partials, int target, hipStream_t s);
hipError_t launch_coarsest_cg(const CoarsestView& C, const bf16* b, bf16* x, hipStream_t s);
hipError_t launch_block_pre(const LevelView& L, bool from_zero, const bf16* x, const bf16* b, bf16* x_out,
                            bf16* coarse_b, float w0, float w1, float w2, hipStream_t s);
hipError_t launch_block_post(const LevelView& L, const bf16* x, const bf16* xc, const bf16* b, bf16* x_out,
                             float w0, float w1, float w2, hipStream_t s);
// ---- FP64 residual replacement: r = b - A x with every product and sum in
// FP64, stored as FP32, then nu = ||r||_inf through the mailbox
hipError_t launch_residual64_absmax(const LevelView& L, const float* x, const float* b, float* r, PcgScalars* sc,
                                    double* partials, HostMailbox* mb, hipStream_t s);
// With FP64 residual replacement, the PCG steps started while ||r||_inf <=
// kReplaceBand * the tolerance form r as the true residual of the new iterate
// (about the last three steps at a tenfold reduction per step), so the
// recursive residual's drift never builds up near convergence.
constexpr float kReplaceBand = 1000.f;
// the same for the iterate with its pending update: r = b - A (x + alpha p),
// alpha from the scalars (the replacement taking the place of r -= alpha A p)
hipError_t launch_replace64_absmax(const LevelView& L, const float* x, const float* p, const float* b, float* r,
                                  PcgScalars* sc, double* partials, HostMailbox* mb, hipStream_t s);

}  // namespace uaamg
