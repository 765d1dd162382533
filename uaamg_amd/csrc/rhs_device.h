// Right-hand side of the pressure Poisson system built on the device
// (LaplacianWithLevel::createPressurePoissonRightHandSide / BuildPoissonRhs,
// simd_vdb_poisson.cpp:19-127): for every level-0 DOF,
//   b = sum over the faces x+, x-, y+, y-, z+, z- of -+(1/dx) (w u + (1 - w) u_solid),
// 0 when every face weight is 0, with the face value of axis c at the voxel
// itself (the - face) or at the voxel + e_c (the + face), terms added in that
// order with the host build's roundings (bit-identical to it).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "poisson_kernels.h"

namespace uaamg {

// One input grid on the device: values [nleaf][512][ncomp] in the grid's own
// leaf order; a missing leaf reads bg[comp] (ValueAccessor::getValue).
struct RhsGridDev {
    const float* values;
    int32_t ncomp;
    float bg[3];
};

enum { kRhsWeight = 0, kRhsVx = 1, kRhsVy = 2, kRhsVz = 3, kRhsSolid = 4, kRhsGrids = 5 };

struct RhsArgs {
    RhsGridDev g[kRhsGrids];
    // [nleaf0][kRhsGrids][4]: leaf of each grid at the DOF leaf's origin and at
    // its +x, +y, +z neighbours' origins (-1: not in the grid)
    const int32_t* tab;
    float invdx;
};

// rhs_dof: ndof floats, DOF order, device memory
hipError_t launch_poisson_rhs(const LevelView& L0, const RhsArgs& a, float* rhs_dof, hipStream_t s);

// A caller's grid: leaves in any order (origins multiples of 8, no
// duplicates), values [nleaf][512][ncomp] in host or device memory.
struct RhsGridRef {
    int64_t nleaf = 0;
    const int32_t* origins = nullptr;
    const float* values = nullptr;
    int32_t ncomp = 1;
    float bg[3] = {0.f, 0.f, 0.f};
    bool values_on_device = false;
};

class PoissonDevice;

// Builds the leaf tables on the host (origin lookups per level-0 leaf),
// uploads host-resident values for the call, and runs the kernel on the
// solver's stream; returns when rhs_dof is written. grids: kRhsGrids entries
// in kRhs* order (an empty solid-velocity grid reads its background).
// Returns 0, or a negative uaamg status with err set.
int build_poisson_rhs_device(PoissonDevice& d, const RhsGridRef* grids, float* rhs_dof, std::string& err);

// Level-0 DOF leaves the allocation-free build takes at most (its leaf tables
// live in a fixed device array of 80 bytes per leaf, 10 MiB).
constexpr int64_t kRhsNoAllocMaxLeaves = 131072;

// The same build with no device allocation: every grid's values must already
// be device memory, and the leaf tables go into the library's fixed per-device
// scratch array, which all calls on one GPU share (calls that overlap on one
// GPU race on it). Returns 0, or a negative uaamg status with err set.
int build_poisson_rhs_device_noalloc(PoissonDevice& d, const RhsGridRef* grids, float* rhs_dof, std::string& err);

}  // namespace uaamg
