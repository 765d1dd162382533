// Face weights from the solid SDF (FLIP_Solver_OpenVDB::calculate_face_weights,
// FLIP_vdb.cpp:1990-2068): the liquid SDF's topology dilated twice over the 26
// neighbours (dilateActiveValues(2, NN_FACE_EDGE_VERTEX)); at each active voxel v
// the weight of face c is 1 - fraction_inside of the solid SDF over the four
// corners v + {0,1}^2 of that face's plane (clamped to [0, 1]); every other
// voxel of the grid holds 1. The topology is built on the host, the weights on
// the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "leaf_topology.h"

namespace uaamg {

// Active voxels of a grid dilated k times over the 26 neighbours (Chebyshev
// distance <= k): leaves in tree order (the input leaves and every leaf a
// dilated voxel lands in) and their masks.
void dilate_active_26(const std::vector<Coord>& origins, const std::vector<uint64_t>& active, int k,
                      std::vector<Coord>& out_origins, std::vector<uint64_t>& out_active);

struct FaceWeightArgs {
    const float* solid;        // solid SDF values [nsolid][512] (device)
    const int32_t* tab;        // [nleaf][8] solid leaf at origin + 8 (i, j, k), i,j,k in {0,1} (k fastest)
    float solid_bg;
    float cut;                 // a clamped weight below this becomes 0 (0: none; the member variant: 1e-2)
};

// values: [nleaf][512][3] (device)
hipError_t launch_face_weights(int64_t nleaf, const uint64_t* active, const FaceWeightArgs& a, float* values,
                               hipStream_t s);

// The whole computation on `device`: topology of the weight grid (origins in
// tree order, masks) and, when values is non-null, the weights [nleaf][512][3]
// in host memory. Returns 0, or a negative uaamg status with err set.
int face_weights_device(const HostGrid& liquid, int64_t nsolid, const int32_t* solid_origins, const float* solid_values,
                        float solid_bg, float cut, int device, std::vector<Coord>& origins,
                        std::vector<uint64_t>& active, float* values, std::string& err);

}  // namespace uaamg
