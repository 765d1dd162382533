// Device side of the viscosity level-0 volume stage (visc_volumes.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace uaamg {

struct VolumeStageArgs {
    const float* liquid;       // liquid SDF values [nliquid][512]
    const int32_t* liq_nb;     // [nsub][27] liquid leaf at sub leaf origin + 8 (ax, ay, az), a in {-1,0,1} (-1: none)
    const int32_t* sub_nb_up;  // [nsub][8] sub leaf at origin + 8 (i, j, k), i,j,k in {0,1} (k fastest)
    const int32_t* vol_sub_nb; // [nvol][27] sub leaf at vol leaf origin + 8 (ax, ay, az)
    float liquid_bg;
    float sdf_bg;              // 1.5 dx: the subcell SDF's clamp and missing-leaf value
};

// sdf, svol: [nsub][4096]; out: 7 arrays of [nvol][512] back to back (voxel, face x/y/z, edge x/y/z)
hipError_t launch_visc_volumes(const VolumeStageArgs& a, int32_t nsub, int32_t nvol, float* sdf, float* svol,
                               float* out, hipStream_t s);

}  // namespace uaamg
