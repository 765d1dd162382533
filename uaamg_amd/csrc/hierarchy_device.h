// The Poisson level hierarchy built on the device, its coefficient bricks
// kept in device memory (LaplacianWithLevel construction,
// simd_vdb_poisson.cpp:260-671 and 1756-1811).
//
// Per level the device forms what the host build forms, with the same
// arithmetic in the same order, so every brick is bit-identical:
//   level 0      the per-voxel stage (finest_device.h) over all liquid-phi
//                leaves, then compaction to the leaves holding DOFs
//                (a gather by the host's list of kept leaves);
//   levels >= 1  the 2^3 Galerkin sums 0.5 * R A P (initializeFromFineLevel,
//                :392-467) from the finer level's trimmed bricks;
//   every level  trimDefaultNodes (:1773-1811): a leaf whose DOF values all
//                lie within |eps| of the default keeps no brick (flag 0) and
//                reads the default; then 1/diag (0 where diag == 0, the
//                default on trimmed leaves and non-DOF voxels).
// The host keeps the topology (origins, DOF masks, numbering, neighbour and
// child tables: bit operations on a few MB) and the keep flags, and downloads
// the coarsest level's bricks for the coarsest solver's explicit matrix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "poisson_host.h"

namespace uaamg {

struct TrimArgs {
    int32_t n;                 // leaves of the level
    const int32_t* gather;     // [n] source leaf of each leaf in src, or null (src == dst, in place)
    const float* src[4];       // diag, x, y, z source bricks
    float* dst[4];             // [n][512] output bricks
    float* invdiag;            // [n][512]
    const uint64_t* mask;      // [n][8] DOF masks of the level
    float dflt[4];             // defaults of the four kinds
    float eps[4];              // |eps| of the four kinds
    float inv_dflt;            // default of 1/diag
    uint8_t* keep;             // out [4][n]: 1 keeps the brick, 0 trimmed
};

struct GalerkinArgs {
    int32_t n;                 // coarse leaves
    const int32_t* child;      // [n][8] fine leaf of octant (bx<<2|by<<1|bz), -1 absent
    const uint64_t* cmask;     // [n][8] coarse DOF masks
    const uint64_t* fmask;     // [nf][8] fine DOF masks
    const int32_t* fnbr;       // [nf][6] fine face neighbours -x,+x,-y,+y,-z,+z
    const float* fine[4];      // [nf][512] trimmed fine bricks
    float* coarse[4];          // out [n][512]
    float term;                // coarse dt / dx^2
};

hipError_t launch_level_trim(const TrimArgs& a, hipStream_t s);
hipError_t launch_galerkin(const GalerkinArgs& a, hipStream_t s);
// dst[slot[i]] = src[i] for every leaf with slot[i] >= 0 (bricks of 512 floats)
hipError_t launch_pack_bricks(int32_t n, const int32_t* slot, const float* src, float* dst, hipStream_t s);

// Builds the hierarchy of liquid phi / face weights on the device: topology
// and keep flags in `out` as build_hierarchy forms them, the bricks of every
// level but the coarsest in out.dcoef (those levels have coef_on_device set
// and empty host arrays). Returns false with err set on a device error.
bool build_hierarchy_device(const HostGrid& phi, const HostGrid& weights, float dt, float dx,
                            int64_t max_coarsest_dof, PoissonHierarchyHost& out, int device, std::string& err);

// Copies the bricks of every level whose coefficients live in h.dcoef into
// the host arrays (coef_on_device cleared). Returns false with err set on a
// device error.
bool download_device_coefficients(PoissonHierarchyHost& h, std::string& err);

}  // namespace uaamg
