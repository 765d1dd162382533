// volume_fraction (volume_fractions.cpp:82-126), shared by host and device code:
// the liquid fraction of a cube from the signed distances at its eight corners,
// as ten tetrahedra (two of them counted twice) over the cube's corners, each
// tetrahedron's fraction from its sorted corner values (tetrahedron and prism
// cases). Operations in the reference's order (built without contraction).
#pragma once

#include <hip/hip_runtime.h>

namespace uaamg {

__host__ __device__ inline float vf_tet_frac(float p0, float p1, float p2, float p3) {
    return p0 * p0 * p0 / ((p0 - p1) * (p0 - p2) * (p0 - p3));
}

__host__ __device__ inline float vf_prism_frac(float p0, float p1, float p2, float p3) {
    const float a = p0 / (p0 - p2), b = p0 / (p0 - p3), c = p1 / (p1 - p3), d = p1 / (p1 - p2);
    return a * b * (1 - d) + b * (1 - c) * d + c * d;
}

__host__ __device__ inline void vf_order(float& a, float& b) {
    if (a > b) {
        const float t = a;
        a = b;
        b = t;
    }
}

__host__ __device__ inline float vf_tet_volume(float p0, float p1, float p2, float p3) {
    // five compare-exchanges sort the corner values ascending
    vf_order(p0, p1);
    vf_order(p2, p3);
    vf_order(p0, p2);
    vf_order(p1, p3);
    vf_order(p1, p2);
    if (p3 <= 0) return 1;
    if (p2 <= 0) return 1 - vf_tet_frac(p3, p2, p1, p0);
    if (p1 <= 0) return vf_prism_frac(p0, p1, p2, p3);
    if (p0 <= 0) return vf_tet_frac(p0, p1, p2, p3);
    return 0;
}

__host__ __device__ inline float volume_fraction8(float p000, float p100, float p010, float p110, float p001,
                                                  float p101, float p011, float p111) {
    if (p000 < 0.f && p100 < 0.f && p010 < 0.f && p110 < 0.f && p001 < 0.f && p101 < 0.f && p011 < 0.f &&
        p111 < 0.f)
        return 1.0f;
    return (vf_tet_volume(p000, p001, p101, p011) + vf_tet_volume(p000, p101, p100, p110) +
            vf_tet_volume(p000, p010, p011, p110) + vf_tet_volume(p101, p011, p111, p110) +
            2 * vf_tet_volume(p000, p011, p101, p110) + vf_tet_volume(p100, p101, p001, p111) +
            vf_tet_volume(p100, p001, p000, p010) + vf_tet_volume(p100, p110, p111, p010) +
            vf_tet_volume(p001, p111, p011, p010) + 2 * vf_tet_volume(p100, p111, p001, p010)) /
           12;
}

}  // namespace uaamg
