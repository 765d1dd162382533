// Device face weights from the solid SDF (face_weights.h).
#include "face_weights.h"
#include "fraction_inside.h"
#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

// std::max(0.f, std::min(v, 1.f)) as the reference writes it (a NaN becomes 0)
__device__ __forceinline__ float clamp01(float v) {
    const float m = (1.f < v) ? 1.f : v;
    return (0.f < m) ? m : 0.f;
}

// one wavefront per leaf of the weight grid, one lane per z-row
__global__ __launch_bounds__(kBlockThreads) void face_weights_kernel(int64_t nleaf, const uint64_t* active,
                                                                     FaceWeightArgs a, float* values) {
    const int leaf = wave_leaf();
    if (leaf >= nleaf) return;
    const int lane = threadIdx.x & 63;
    const int lx = lane >> 3, ly = lane & 7;
    const uint32_t rm = uint32_t(active[size_t(leaf) * 8 + lx] >> (ly * 8)) & 0xffu;
    const int32_t* t = a.tab + size_t(leaf) * 8;
    for (int z = 0; z < 8; ++z) {
        float* out = values + (size_t(leaf) * 512 + size_t((lx << 6) | (ly << 3) | z)) * 3;
        if (!((rm >> z) & 1u)) {
            out[0] = out[1] = out[2] = 1.0f;
            continue;
        }
        // solid SDF at the eight corners v + (i, j, k)
        float s[2][2][2];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) {
                    int X = lx + i, Y = ly + j, Z = z + k;
                    const int nb = ((X >> 3) << 2) | ((Y >> 3) << 1) | (Z >> 3);
                    X &= 7;
                    Y &= 7;
                    Z &= 7;
                    const int32_t sl = t[nb];
                    s[i][j][k] = sl >= 0 ? a.solid[size_t(sl) * 512 + size_t((X << 6) | (Y << 3) | Z)] : a.solid_bg;
                }
        const float fu = fraction_inside4(s[0][0][0], s[0][1][0], s[0][0][1], s[0][1][1]);
        const float fv = fraction_inside4(s[0][0][0], s[0][0][1], s[1][0][0], s[1][0][1]);
        const float fw = fraction_inside4(s[0][0][0], s[1][0][0], s[0][1][0], s[1][1][0]);
        float w[3] = {clamp01(1.0f - fu), clamp01(1.0f - fv), clamp01(1.0f - fw)};
        for (int c = 0; c < 3; ++c)
            if (w[c] < a.cut) w[c] = 0.f;
        out[0] = w[0];
        out[1] = w[1];
        out[2] = w[2];
    }
}

}  // namespace

hipError_t launch_face_weights(int64_t nleaf, const uint64_t* active, const FaceWeightArgs& a, float* values,
                               hipStream_t s) {
    if (nleaf <= 0) return hipSuccess;
    if (nleaf > int64_t(INT32_MAX) / 64) return hipErrorInvalidValue;
    face_weights_kernel<<<grid_for(int32_t(nleaf)), kBlockThreads, 0, s>>>(nleaf, active, a, values);
    return hipGetLastError();
}

}  // namespace uaamg
