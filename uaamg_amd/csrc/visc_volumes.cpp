// Host side of the viscosity level-0 volume stage on the device (visc_volumes.h).
#include "visc_volumes.h"

#include <hip/hip_runtime.h>

#include "visc_volumes_kernels.h"

namespace uaamg {

namespace {

struct DevMem {
    void* p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

// leaf index of `lk` at origin + 8 * offset for each of the 27 (or 8) offsets of every origin
std::vector<int32_t> neighbour_table(const std::vector<Coord>& org, const LeafLookup& lk, bool up_only) {
    const int m = up_only ? 8 : 27;
    std::vector<int32_t> t(org.size() * size_t(m));
    for (size_t i = 0; i < org.size(); ++i)
        for (int d = 0; d < m; ++d) {
            const int ax = up_only ? (d >> 2) : d / 9 - 1, ay = up_only ? ((d >> 1) & 1) : (d / 3) % 3 - 1,
                      az = up_only ? (d & 1) : d % 3 - 1;
            t[i * size_t(m) + size_t(d)] = lk.find({org[i].x + 8 * ax, org[i].y + 8 * ay, org[i].z + 8 * az});
        }
    return t;
}

}  // namespace

int visc_volumes_device(int device, const HostGrid& liquid, const std::vector<Coord>& sub_origin,
                        const std::vector<Coord>& vol_origin, float dx, std::vector<float>& voxel_vol,
                        std::vector<float> face_vol[3], std::vector<float> edge_vol[3], std::string& err) {
    const size_t nsub = sub_origin.size(), nvol = vol_origin.size(), nliq = liquid.nleaf();
    if (nsub > size_t(INT32_MAX) || nvol > size_t(INT32_MAX)) {
        err = "volume stage: too many leaves";
        return -1;
    }
    if (hipSetDevice(device) != hipSuccess) {
        err = "hipSetDevice failed";
        return -4;
    }
    LeafLookup sub_lk;
    sub_lk.build(sub_origin);
    const std::vector<int32_t> liq_nb = neighbour_table(sub_origin, liquid.lookup, false);
    const std::vector<int32_t> sub_up = neighbour_table(sub_origin, sub_lk, true);
    const std::vector<int32_t> vol_nb = neighbour_table(vol_origin, sub_lk, false);

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
        err = std::string("volume stage, ") + what + ": " + hipGetErrorString(e);
        return -3;
    };
    DevMem liq_d, liq_nb_d, sub_up_d, vol_nb_d, sdf_d, svol_d, out_d;
    auto upload = [&](DevMem& m, const void* src, size_t bytes) -> hipError_t {
        if (hipError_t e = hipMalloc(&m.p, bytes > 0 ? bytes : 4)) return e;
        return bytes ? hipMemcpyAsync(m.p, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
    };
    if (hipError_t e = upload(liq_d, liquid.values, nliq * kLeafSize * sizeof(float))) return hip_fail(e, "liquid");
    if (hipError_t e = upload(liq_nb_d, liq_nb.data(), liq_nb.size() * sizeof(int32_t))) return hip_fail(e, "tables");
    if (hipError_t e = upload(sub_up_d, sub_up.data(), sub_up.size() * sizeof(int32_t))) return hip_fail(e, "tables");
    if (hipError_t e = upload(vol_nb_d, vol_nb.data(), vol_nb.size() * sizeof(int32_t))) return hip_fail(e, "tables");
    const size_t sub_bytes = nsub * 4096 * sizeof(float), out_n = nvol * kLeafSize;
    if (hipError_t e = hipMalloc(&sdf_d.p, sub_bytes > 0 ? sub_bytes : 4)) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&svol_d.p, sub_bytes > 0 ? sub_bytes : 4)) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&out_d.p, out_n > 0 ? 7 * out_n * sizeof(float) : 4)) return hip_fail(e, "hipMalloc");
    VolumeStageArgs a{};
    a.liquid = static_cast<const float*>(liq_d.p);
    a.liq_nb = static_cast<const int32_t*>(liq_nb_d.p);
    a.sub_nb_up = static_cast<const int32_t*>(sub_up_d.p);
    a.vol_sub_nb = static_cast<const int32_t*>(vol_nb_d.p);
    a.liquid_bg = liquid.background[0];
    a.sdf_bg = 1.5f * dx;
    if (hipError_t e = launch_visc_volumes(a, int32_t(nsub), int32_t(nvol), static_cast<float*>(sdf_d.p),
                                           static_cast<float*>(svol_d.p), static_cast<float*>(out_d.p), s))
        return hip_fail(e, "kernels");
    voxel_vol.assign(out_n, 0.f);
    for (int c = 0; c < 3; ++c) {
        face_vol[c].assign(out_n, 0.f);
        edge_vol[c].assign(out_n, 0.f);
    }
    if (out_n) {
        const float* o = static_cast<const float*>(out_d.p);
        std::vector<float>* dst[7] = {&voxel_vol, &face_vol[0], &face_vol[1], &face_vol[2],
                                      &edge_vol[0], &edge_vol[1], &edge_vol[2]};
        for (int k = 0; k < 7; ++k)
            if (hipError_t e = hipMemcpyAsync(dst[k]->data(), o + size_t(k) * out_n, out_n * sizeof(float),
                                              hipMemcpyDeviceToHost, s))
                return hip_fail(e, "download");
    }
    if (hipError_t e = hipStreamSynchronize(s)) return hip_fail(e, "sync");
    return 0;
}

}  // namespace uaamg
