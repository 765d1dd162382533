// Host side of the face-weight computation (face_weights.h).
#include "face_weights.h"

#include <algorithm>
#include <unordered_map>

#include "poisson_host.h"

namespace uaamg {

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

void dilate_active_26(const std::vector<Coord>& origins, const std::vector<uint64_t>& active, int k,
                      std::vector<Coord>& out_origins, std::vector<uint64_t>& out_active) {
    // candidate leaves: the inputs and their 26 neighbours (k <= 8 keeps a
    // dilated voxel within one leaf of its source)
    std::unordered_map<uint64_t, int32_t> index;
    std::vector<Coord> cand;
    auto add = [&](const Coord& o) {
        if (index.emplace(pack_origin(o), int32_t(cand.size())).second) cand.push_back(o);
    };
    for (const Coord& o : origins) add(o);
    for (const Coord& o : origins)
        for (int d = 0; d < 27; ++d) add({o.x + 8 * (d / 9 - 1), o.y + 8 * ((d / 3) % 3 - 1), o.z + 8 * (d % 3 - 1)});
    LeafLookup in;
    in.build(origins);
    std::vector<uint64_t> mask(cand.size() * 8, 0);
    parallel_for(int64_t(cand.size()), [&](int64_t b, int64_t e) {
        // the 24^3 neighbourhood of a leaf as 24x24 rows of 24 z-bits
        std::vector<uint32_t> vol(24 * 24), tmp(24 * 24);
        for (int64_t ci = b; ci < e; ++ci) {
            const Coord o = cand[size_t(ci)];
            std::fill(vol.begin(), vol.end(), 0u);
            bool any = false;
            for (int d = 0; d < 27; ++d) {
                const int ax = d / 9, ay = (d / 3) % 3, az = d % 3;
                const int32_t li = in.find({o.x + 8 * (ax - 1), o.y + 8 * (ay - 1), o.z + 8 * (az - 1)});
                if (li < 0) continue;
                const uint64_t* m = &active[size_t(li) * 8];
                for (int x = 0; x < 8; ++x)
                    for (int y = 0; y < 8; ++y) {
                        const uint32_t row = uint32_t(m[x] >> (y * 8)) & 0xffu;
                        if (!row) continue;
                        any = true;
                        vol[size_t((ax * 8 + x) * 24 + ay * 8 + y)] |= row << (az * 8);
                    }
            }
            if (!any) continue;
            // separable box dilation of radius k: z (bits), then y, then x
            for (auto& r : vol) {
                uint32_t d = r;
                for (int s = 1; s <= k; ++s) d |= (r << s) | (r >> s);
                r = d;
            }
            for (int x = 0; x < 24; ++x)
                for (int y = 0; y < 24; ++y) {
                    uint32_t d = 0;
                    for (int s = -k; s <= k; ++s)
                        if (y + s >= 0 && y + s < 24) d |= vol[size_t(x * 24 + y + s)];
                    tmp[size_t(x * 24 + y)] = d;
                }
            for (int x = 0; x < 24; ++x)
                for (int y = 0; y < 24; ++y) {
                    uint32_t d = 0;
                    for (int s = -k; s <= k; ++s)
                        if (x + s >= 0 && x + s < 24) d |= tmp[size_t((x + s) * 24 + y)];
                    vol[size_t(x * 24 + y)] = d;
                }
            uint64_t* out = &mask[size_t(ci) * 8];
            for (int x = 0; x < 8; ++x)
                for (int y = 0; y < 8; ++y)
                    out[x] |= uint64_t((vol[size_t((8 + x) * 24 + 8 + y)] >> 8) & 0xffu) << (y * 8);
        }
    });
    // keep the input leaves (even empty) and every leaf a dilated voxel lands in, in tree order
    std::vector<size_t> keep;
    for (size_t i = 0; i < cand.size(); ++i) {
        bool on = i < origins.size();
        for (int w = 0; w < 8 && !on; ++w) on = mask[i * 8 + size_t(w)] != 0;
        if (on) keep.push_back(i);
    }
    std::sort(keep.begin(), keep.end(), [&](size_t a, size_t b) { return tree_order_less(cand[a], cand[b]); });
    out_origins.resize(keep.size());
    out_active.resize(keep.size() * 8);
    for (size_t j = 0; j < keep.size(); ++j) {
        out_origins[j] = cand[keep[j]];
        std::copy(&mask[keep[j] * 8], &mask[keep[j] * 8] + 8, &out_active[j * 8]);
    }
}

int face_weights_device(const HostGrid& liquid, int64_t nsolid, const int32_t* solid_origins, const float* solid_values,
                        float solid_bg, float cut, int device, std::vector<Coord>& origins,
                        std::vector<uint64_t>& active, float* values, std::string& err) {
    dilate_active_26(liquid.origin, liquid.active, 2, origins, active);
    if (!values) return 0;
    if (hipSetDevice(device) != hipSuccess) {
        err = "hipSetDevice failed";
        return -4;
    }
    hipStream_t s = nullptr;
    if (hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) {
        err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        return -3;
    }
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    } guard{s};
    auto hip_fail = [&](hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return -3;
    };
    std::vector<Coord> so(static_cast<size_t>(nsolid));
    for (int64_t i = 0; i < nsolid; ++i) {
        so[size_t(i)] = {solid_origins[3 * i], solid_origins[3 * i + 1], solid_origins[3 * i + 2]};
        if ((so[size_t(i)].x | so[size_t(i)].y | so[size_t(i)].z) & 7) {
            err = "solid_sdf: leaf origin not a multiple of 8";
            return -1;
        }
    }
    LeafLookup sl;
    sl.build(so);
    for (size_t i = 0; i < so.size(); ++i)
        if (sl.find(so[i]) != int32_t(i)) {
            err = "solid_sdf: duplicate leaf origins";
            return -1;
        }
    const size_t n = origins.size();
    std::vector<int32_t> tab(n * 8);
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 8; ++d)
            tab[i * 8 + size_t(d)] =
                sl.find({origins[i].x + 8 * (d >> 2), origins[i].y + 8 * ((d >> 1) & 1), origins[i].z + 8 * (d & 1)});
    DevBuf solid_d, tab_d, act_d, val_d;
    const size_t sbytes = size_t(nsolid) * kLeafSize * sizeof(float);
    const size_t vbytes = n * kLeafSize * 3 * sizeof(float);
    if (nsolid > 0) {
        if (hipError_t e = hipMalloc(&solid_d.p, sbytes)) return hip_fail(e, "hipMalloc");
        if (hipError_t e = hipMemcpyAsync(solid_d.p, solid_values, sbytes, hipMemcpyHostToDevice, s))
            return hip_fail(e, "upload solid sdf");
    }
    if (n == 0) return 0;
    if (hipError_t e = hipMalloc(&tab_d.p, tab.size() * sizeof(int32_t))) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&act_d.p, active.size() * sizeof(uint64_t))) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&val_d.p, vbytes)) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMemcpyAsync(tab_d.p, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice, s))
        return hip_fail(e, "upload tables");
    if (hipError_t e =
            hipMemcpyAsync(act_d.p, active.data(), active.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s))
        return hip_fail(e, "upload masks");
    FaceWeightArgs a{static_cast<const float*>(solid_d.p), static_cast<const int32_t*>(tab_d.p), solid_bg, cut};
    if (hipError_t e = launch_face_weights(int64_t(n), static_cast<const uint64_t*>(act_d.p), a,
                                           static_cast<float*>(val_d.p), s))
        return hip_fail(e, "face weights kernel");
    if (hipError_t e = hipMemcpyAsync(values, val_d.p, vbytes, hipMemcpyDeviceToHost, s))
        return hip_fail(e, "download weights");
    if (hipError_t e = hipStreamSynchronize(s)) return hip_fail(e, "face weights sync");
    return 0;
}

}  // namespace uaamg
