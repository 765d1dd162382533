// fraction_inside (levelset_util.cpp:4-103), shared by host and device code:
// the fraction of a segment / square that lies inside (phi < 0), from the
// signed distances at its ends / corners.
#pragma once

#include <hip/hip_runtime.h>

namespace uaamg {

__host__ __device__ inline float fraction_inside2(float phi_left, float phi_right) {
    if (phi_left < 0 && phi_right < 0) return 1;
    if (phi_left < 0 && phi_right >= 0) return phi_left / (phi_left - phi_right);
    if (phi_left >= 0 && phi_right < 0) return phi_right / (phi_right - phi_left);
    return 0;
}

__host__ __device__ inline void rotate4(float* a) {
    const float t = a[0];
    a[0] = a[1];
    a[1] = a[2];
    a[2] = a[3];
    a[3] = t;
}

// corners bottom-left, bottom-right, top-left, top-right
__host__ __device__ inline float fraction_inside4(float phi_bl, float phi_br, float phi_tl, float phi_tr) {
    const int inside = (phi_bl < 0) + (phi_tl < 0) + (phi_br < 0) + (phi_tr < 0);
    float q[4] = {phi_bl, phi_br, phi_tr, phi_tl};   // counter-clockwise corner list
    if (inside == 4) return 1;
    if (inside == 3) {
        while (q[0] < 0) rotate4(q);   // the single outside corner goes first
        const float s0 = 1 - fraction_inside2(q[0], q[3]);
        const float s1 = 1 - fraction_inside2(q[0], q[1]);
        return 1 - 0.5f * s0 * s1;
    }
    if (inside == 2) {
        while (q[0] >= 0 || !(q[1] < 0 || q[2] < 0)) rotate4(q);
        if (q[1] < 0) {   // the two inside corners share an edge
            const float left = fraction_inside2(q[0], q[3]);
            const float right = fraction_inside2(q[1], q[2]);
            return 0.5f * (left + right);
        }
        // diagonal pair: the centre value picks the topology
        const float mid = 0.25f * (q[0] + q[1] + q[2] + q[3]);
        if (mid < 0) {
            float area = 0;
            const float s1 = 1 - fraction_inside2(q[0], q[3]);
            const float s3 = 1 - fraction_inside2(q[2], q[3]);
            area += 0.5f * s1 * s3;
            const float s2 = 1 - fraction_inside2(q[2], q[1]);
            const float s0 = 1 - fraction_inside2(q[0], q[1]);
            area += 0.5f * s0 * s2;
            return 1 - area;
        }
        float area = 0;
        const float s0 = fraction_inside2(q[0], q[1]);
        const float s1 = fraction_inside2(q[0], q[3]);
        area += 0.5f * s0 * s1;
        const float s2 = fraction_inside2(q[2], q[1]);
        const float s3 = fraction_inside2(q[2], q[3]);
        area += 0.5f * s2 * s3;
        return area;
    }
    if (inside == 1) {
        while (q[0] >= 0) rotate4(q);
        const float s0 = fraction_inside2(q[0], q[3]);
        const float s1 = fraction_inside2(q[0], q[1]);
        return 0.5f * s0 * s1;
    }
    return 0;
}

}  // namespace uaamg
