// Wavefront reductions on DPP (data-parallel primitives): register-to-register
// lane exchanges that the VALU issues back to back, instead of the LDS
// crossbar round trips of a shuffle butterfly.
#pragma once

#include <hip/hip_runtime.h>

namespace uaamg {
namespace dev {

template <int kCtrl>
__device__ __forceinline__ float dpp_move(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, 0xf, 0xf, false));
}

__device__ __forceinline__ float lane_value(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// Sum of a wavefront's 64 values, the same in every lane and independent of
// the lane: rotations by 8, 4, 2, 1 within each 16-lane row (row_ror) leave
// the row's sum in its lanes; the four rows' lane-0 sums are then added in a
// fixed order.
__device__ __forceinline__ float wave_sum_dpp(float v) {
    v += dpp_move<0x128>(v);
    v += dpp_move<0x124>(v);
    v += dpp_move<0x122>(v);
    v += dpp_move<0x121>(v);
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

}  // namespace dev
}  // namespace uaamg
