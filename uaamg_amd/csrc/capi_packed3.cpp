// packed_FloatGrid3's conversions between a Vec3f grid and three single-channel
// grids (packed3grids.cpp:11-83), on flattened leaves. These are host copies:
// a caller keeps its velocity either way, and the solver reads the channels
// through uaamg_viscosity_build_rhs_packed / uaamg_viscosity_grid_to_vector.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "capi_common.h"

using namespace uaamg;
using namespace uaamg::capi;

extern "C" {

int uaamg_packed3_from_vec3(const uaamg_grid* vec3, float* values) {
    UAAMG_GUARD_BEGIN
    if (!vec3 || !values) return fail(UAAMG_EINVAL, "null argument");
    if (vec3->ncomp != 3) return fail(UAAMG_EINVAL, "vec3: expected a 3-component grid");
    const int64_t n = vec3->nleaf;
    if (n < 0 || (n > 0 && (!vec3->origins || !vec3->values)))
        return fail(UAAMG_EINVAL, "vec3: null arrays");
    // each channel: the Vec3 grid's topology with every value at the channel's
    // background (a topology copy), then the channel of every active voxel
    for (int64_t i = 0; i < n; ++i) {
        const float* src = vec3->values + size_t(i) * kLeafSize * 3;
        const uint64_t* m = vec3->active ? vec3->active + size_t(i) * 8 : nullptr;
        for (int c = 0; c < 3; ++c) {
            float* dst = values + (size_t(c) * size_t(n) + size_t(i)) * kLeafSize;
            const float bg = vec3->background[c];
            for (int off = 0; off < kLeafSize; ++off)
                dst[off] = (!m || mask_test(m, off)) ? src[size_t(off) * 3 + size_t(c)] : bg;
        }
    }
    return 0;
    UAAMG_GUARD_END
}

int uaamg_packed3_to_vec3(const uaamg_grid* v0, const uaamg_grid* v1, const uaamg_grid* v2, int64_t* nleaf,
                          int32_t* origins, uint64_t* active, float* values) {
    UAAMG_GUARD_BEGIN
    if (!v0 || !v1 || !v2 || !nleaf) return fail(UAAMG_EINVAL, "null argument");
    const uaamg_grid* in[3] = {v0, v1, v2};
    HostGrid g[3];
    for (int c = 0; c < 3; ++c)
        if (!to_host_grid(in[c], 1, g[c], "channel")) return UAAMG_EINVAL;
    // the union of the three topologies, in tree order
    std::vector<Coord> org;
    for (int c = 0; c < 3; ++c) org.insert(org.end(), g[c].origin.begin(), g[c].origin.end());
    std::sort(org.begin(), org.end(), tree_order_less);
    org.erase(std::unique(org.begin(), org.end()), org.end());
    const int64_t n = int64_t(org.size());
    if (!origins && !active && !values) {
        *nleaf = n;
        return 0;
    }
    if (*nleaf < n) {
        *nleaf = n;
        return fail(UAAMG_EINVAL, "output arrays hold fewer leaves than the union topology");
    }
    *nleaf = n;
    for (int64_t i = 0; i < n; ++i) {
        const Coord o = org[size_t(i)];
        if (origins) {
            origins[3 * i] = o.x;
            origins[3 * i + 1] = o.y;
            origins[3 * i + 2] = o.z;
        }
        uint64_t m[8] = {};
        float* dst = values ? values + size_t(i) * kLeafSize * 3 : nullptr;
        if (dst) std::fill(dst, dst + kLeafSize * 3, 0.f);   // the new tree's background (0, 0, 0)
        for (int c = 0; c < 3; ++c) {
            const int32_t gi = g[c].lookup.find(o);
            if (gi < 0) continue;
            const uint64_t* cm = &g[c].active[size_t(gi) * 8];
            const float* src = g[c].leaf_values(size_t(gi));
            for (int w = 0; w < 8; ++w) m[w] |= cm[w];
            if (dst)
                for (int off = 0; off < kLeafSize; ++off)
                    if (mask_test(cm, off)) dst[size_t(off) * 3 + size_t(c)] = src[off];
        }
        if (active) std::memcpy(active + size_t(i) * 8, m, sizeof(m));
    }
    return 0;
    UAAMG_GUARD_END
}

}  // extern "C"
