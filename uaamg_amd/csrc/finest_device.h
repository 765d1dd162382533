// The per-voxel stage of the level-0 Poisson build on the device
// (LaplacianWithLevel::BuildFinestMatrix / initializeFinest,
// simd_vdb_poisson.cpp:474-635): for every voxel of every liquid-phi leaf, the
// DOF test (active, phi < 0, nonzero diagonal), the diagonal as the sum over the
// faces +x,-x,+y,-y,+z,-z of w dt/dx^2 (liquid neighbour) or w dt/dx^2 / theta
// (air neighbour, theta = fraction_inside clamped to >= 0.02), and the -x/-y/-z
// couplings -w dt/dx^2 to liquid neighbours; non-DOF voxels keep the defaults
// 6 dt/dx^2 and -dt/dx^2. Same operations and order as the host build
// (bit-identical); compaction, DOF numbering and trimming stay on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "leaf_topology.h"

namespace uaamg {

struct FinestArgs {
    const float* phi;          // [nphi][512]
    const uint64_t* pact;      // [nphi][8]
    const int32_t* pnb;        // [nphi][6] phi leaf at -x,+x,-y,+y,-z,+z (-1: background)
    const float* w;            // [nw][512][3]
    const int32_t* wtab;       // [nphi][4] weight leaf at the phi leaf's origin and its +x, +y, +z neighbours
    float phi_bg, w_bg[3], term;
    uint64_t* dmask;           // out [nphi][8]
    float* cd;                 // out [nphi][512]
    float* cx;
    float* cy;
    float* cz;
};

hipError_t launch_finest(int64_t nphi, const FinestArgs& a, hipStream_t s);

// Leaf tables of the stage: pnb [nphi][6] as FinestArgs::pnb, wtab [nphi][4]
// as FinestArgs::wtab.
void finest_tables(const HostGrid& phi, const HostGrid& weights, std::vector<int32_t>& pnb,
                   std::vector<int32_t>& wtab);

// phi, weights: tree-ordered host grids. Outputs (caller arrays, every element
// written): dmask [nphi][8], cd and cxyz[0..2] [nphi][512]. Returns 0, or a
// negative uaamg status with err set.
int finest_coefficients_device(const HostGrid& phi, const HostGrid& weights, float term, int device, uint64_t* dmask,
                               float* cd, float* const cxyz[3], std::string& err);

}  // namespace uaamg
