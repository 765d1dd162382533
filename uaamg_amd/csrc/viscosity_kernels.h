// Device kernels of the variational viscosity solve on 8^3 leaf bricks (gfx950).
//
// Layout in HBM, per level: a vector is [leaf][channel][512] floats (the three
// face-velocity channels of a leaf side by side, voxel offset 64x+8y+z) over
// the leaves holding a DOF of any channel; entries that are not DOFs hold 0.
// Coefficient grids are 512-float bricks in one pool, found through
// slot[leaf][grid] (-1: the level's default constant). One wavefront works
// on one (leaf, channel) pair, one lane on one z-row (x,y) = (lane>>3, lane&7).
// Every neighbour value is "the value at a global voxel": the leaf is found in
// a 3x3x3 leaf-neighbourhood table, a missing leaf reads 0 (vectors) or the
// default (coefficients), which is how the reference's applier reads absent
// leaves (simd_viscosity3d.cpp:2991-3000, vdb_SIMD_IO.h:5-191).
#pragma once

#include <algorithm>
#include <cstdlib>

#include <hip/hip_runtime.h>

#include <cstdint>

#include "poisson_kernels.h"   // bf16

namespace uaamg {

// Wave record: one per entry of chan_leaves (a (channel c, DOF leaf) pair,
// one wavefront's work), 64 int32 = 256 bytes, loaded as one burst of scalar
// loads so that every data load of the stencil row depends on one memory
// round trip:
//   [kVRecLeaf]      the leaf
//   [kVRecNb + k]    its 27-leaf neighbourhood, k = (ax+1)*9 + (ay+1)*3 + (az+1) (-1 none; centre = leaf)
//   [kVRecOwn + i]   the leaf's slots of the grids channel c reads: neg(c, 0..2), cross(d2, 0..3),
//                    cross(d1, 0..3), diag c, invdiag c, SPAI0 c   (d1 = c+1, d2 = c+2 mod 3)
//   [kVRecNbr + j]   slots read from a neighbour leaf: neg(c,0) of +x, neg(c,1) of +y, neg(c,2) of +z,
//                    cross(d2,2), cross(d2,3), cross(d1,1), cross(d1,3) of the -e_c neighbour
// (-1: leaf or brick absent, the level default is read).
constexpr int kVRec = 64;
enum { kVRecLeaf = 0, kVRecNb = 1, kVRecOwn = 28, kVRecNbr = 42 };

struct ViscLevelView {
    int32_t nleaf;             // leaves of the table (DOF leaves + coefficient-only leaves)
    int32_t nleaf_dof;         // leaves holding vector bricks
    const int32_t* nb27;       // [nleaf][27]
    const uint64_t* mask;      // [nleaf_dof][3][8]
    const int32_t* slot;       // [nleaf][32]
    const float* pool;         // [nslot][512]
    const int32_t* chan_leaves;   // DOF leaves of channel 0, then 1, then 2
    const int32_t* wrec;          // [chan_off[3]][kVRec] wave records, chan_leaves order
    int32_t chan_off[4];
    const int32_t* parent;     // [nleaf_dof] coarse leaf * 8 + octant
    const int32_t* child;      // [nleaf_dof][8]
    float dflt[32];
    // 1: the colour passes and appliers launch more waves than one wave per
    // SIMD holds at once; their kernels then keep two waves per SIMD
    // resident (at most 256 registers each) rather than all of one row's
    // loads in flight in one wave
    int32_t wide;
    // 1: the colour passes and appliers run on half rows (two wavefronts per
    // wave record, z 0-3 and 4-7), each wavefront with half the operands in
    // flight; 2: the colour passes on quarter rows (four wavefronts); the
    // same values
    int32_t half;
};

// Whether a level's passes take the two-waves-per-SIMD kernels: when a
// channel's pass launches more waves (one per leaf) than one wave per SIMD
// of the device holds at once (256 CUs x 4 SIMDs). UAAMG_VWIDE=0 / =1 forces
// the choice for every level (A/B timing; results are the same either way).
inline int32_t visc_wide(const ViscLevelView& v) {
    if (const char* env = std::getenv("UAAMG_VWIDE")) return env[0] == '1' ? 1 : 0;
    int32_t widest = 0;
    for (int c = 0; c < 3; ++c) widest = std::max(widest, v.chan_off[c + 1] - v.chan_off[c]);
    return widest > 1024 ? 1 : 0;
}

// Whether a level's passes run on part rows: half rows on the levels whose
// passes fit in one wave per SIMD (the ones visc_wide leaves alone). There a
// pass is bound by each wavefront's chain of memory round trips, which
// shorter rows shorten (128^3, us per pass back to back, full / half /
// quarter rows: level 1 (356 leaves) 5.9 / 5.1 / 6.2, level 2 (56) 5.8 / 5.1
// / 4.8, level 3 (12) 5.1 / 5.0 / 4.9; the solve: 13.1 / 12.5 ms, and 12.6 ms
// with quarter rows on levels 2 and 3); on wider levels the extra wavefronts
// cost more (level 0: 9.9 / 12.6 / 16.9). UAAMG_VHALF=0 / 1 / 2 forces full
// / half / quarter rows on every level (results are the same either way).
inline int32_t visc_half(const ViscLevelView& v) {
    if (const char* env = std::getenv("UAAMG_VHALF")) return env[0] == '2' ? 2 : env[0] == '1' ? 1 : 0;
    int32_t widest = 0;
    for (int c = 0; c < 3; ++c) widest = std::max(widest, v.chan_off[c + 1] - v.chan_off[c]);
    return widest <= 1024 ? 1 : 0;
}

// applier modes (light_weight_applier working_mode, simd_viscosity3d.h:13)
enum ViscMode : int { kVNormal = 0, kVResidual = 1, kVJacobi = 2, kVSpai0Mode = 3, kVRedGS = 4, kVBlackGS = 5 };

// Eigen-CG restatement of the coarsest level on its explicit rows
struct ViscCoarsestView {
    int32_t n;
    const int64_t* ptr;        // [n+1]
    const int32_t* col;        // column ascending
    const float* val;
    const float* invdiag;      // Eigen DiagonalPreconditioner: 1/diag or 1
    const int32_t* pos;        // [n] linear index of the DOF in the level's vector
    // the same rows as ELL, entry k of row d at [k * n + d] (columns ascending,
    // rows padded to kViscMaxRow with value 0 and column d)
    const int32_t* ecol;
    const float* eval;
    const int32_t* elen;       // [n] entries of each row
    int32_t ewidth;            // longest row, <= kViscMaxRow
    int32_t max_iters;
    float tol;
};

constexpr int kViscMaxCoarsest = 4096;
constexpr int kViscMaxRow = 16;    // 1 + 6 + 8 entries per row at most

// Level vectors are FP32, or BF16 on the coarse levels of the mixed form
// (precision 1): loads widen exactly, stores round to nearest even, and every
// operation computes in FP32 as in the FP32 form.

// y = op(x, b) on channel c of a level (Gauss-Seidel modes in place: y == x)
hipError_t launch_visc_apply(const ViscLevelView& L, int mode, int channel, const float* x, const float* b, float* y,
                             hipStream_t s);
hipError_t launch_visc_apply(const ViscLevelView& L, int mode, int channel, const bf16* x, const bf16* b, bf16* y,
                             hipStream_t s);
// y = op(x, b) on all three channels in one launch, for the modes with y != x
// (NORMAL, RESIDUAL, JACOBI, SPAI0)
hipError_t launch_visc_apply_all(const ViscLevelView& L, int mode, const float* x, const float* b, float* y,
                                 hipStream_t s);
hipError_t launch_visc_apply_all(const ViscLevelView& L, int mode, const bf16* x, const bf16* b, bf16* y,
                                 hipStream_t s);
// coarse = restriction of fine (restriction, simd_viscosity3d.cpp:2353-2413), all channels
hipError_t launch_visc_restrict(const ViscLevelView& coarse, const ViscLevelView& fine, const float* fine_v,
                                float* coarse_v, hipStream_t s);
hipError_t launch_visc_restrict(const ViscLevelView& coarse, const ViscLevelView& fine, const float* fine_v,
                                bf16* coarse_v, hipStream_t s);
hipError_t launch_visc_restrict(const ViscLevelView& coarse, const ViscLevelView& fine, const bf16* fine_v,
                                bf16* coarse_v, hipStream_t s);
// fine += coarse on fine DOFs (prolongation<true>, :2659-2729), all channels
hipError_t launch_visc_prolong(const ViscLevelView& fine, const ViscLevelView& coarse, const float* coarse_v,
                               float* fine_v, hipStream_t s);
hipError_t launch_visc_prolong(const ViscLevelView& fine, const ViscLevelView& coarse, const bf16* coarse_v,
                               float* fine_v, hipStream_t s);
hipError_t launch_visc_prolong(const ViscLevelView& fine, const ViscLevelView& coarse, const bf16* coarse_v,
                               bf16* fine_v, hipStream_t s);
// x = coarsest solve of b (vectors in level layout), one workgroup
hipError_t launch_visc_coarsest(const ViscCoarsestView& C, const float* b, float* x, hipStream_t s);
hipError_t launch_visc_coarsest(const ViscCoarsestView& C, const bf16* b, bf16* x, hipStream_t s);
// `count` systems in one launch, one workgroup each (views: device array); max_n its largest row count
hipError_t launch_visc_bcoarsest(const ViscCoarsestView* views, int count, int max_n, const float* b, float* x,
                                 hipStream_t s);
// zero every brick of a level vector (set_grid_to_zero, :2304-2315)
hipError_t launch_visc_zero(const ViscLevelView& L, float* v, hipStream_t s);
hipError_t launch_visc_zero(const ViscLevelView& L, bf16* v, hipStream_t s);

// One workgroup per half-visit of a small level (same passes, one launch):
// pre = (zero x when precon,) 2 XYZ sweeps, residual into t, restriction into
// coarse_b; post = x += P x_coarse, 2 ZYX sweeps.
hipError_t launch_visc_small_pre(const ViscLevelView& L, const ViscLevelView& coarse, float* x, const float* b,
                                 float* t, float* coarse_b, bool precon, hipStream_t s);
hipError_t launch_visc_small_pre(const ViscLevelView& L, const ViscLevelView& coarse, bf16* x, const bf16* b,
                                 bf16* t, bf16* coarse_b, bool precon, hipStream_t s);
hipError_t launch_visc_small_post(const ViscLevelView& L, float* x, const float* b, const float* xc, hipStream_t s);
hipError_t launch_visc_small_post(const ViscLevelView& L, bf16* x, const bf16* b, const bf16* xc, hipStream_t s);

}  // namespace uaamg
