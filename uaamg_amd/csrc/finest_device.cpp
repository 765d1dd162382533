// Host side of the device per-voxel level-0 build (finest_device.h).
#include "finest_device.h"

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

void finest_tables(const HostGrid& phi, const HostGrid& weights, std::vector<int32_t>& pnb,
                   std::vector<int32_t>& wtab) {
    const size_t n = phi.nleaf();
    pnb.resize(n * 6);
    wtab.resize(n * 4);
    parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
        for (int64_t li = b; li < e; ++li) {
            const size_t i = size_t(li);
            const Coord o = phi.origin[i];
            for (int c = 0; c < 3; ++c) {
                Coord lo = o, hi = o;
                lo[c] -= 8;
                hi[c] += 8;
                pnb[i * 6 + size_t(2 * c)] = phi.lookup.find(lo);
                pnb[i * 6 + size_t(2 * c + 1)] = phi.lookup.find(hi);
                wtab[i * 4 + size_t(1 + c)] = weights.lookup.find(hi);
            }
            wtab[i * 4] = weights.lookup.find(o);
        }
    });
}

int finest_coefficients_device(const HostGrid& phi, const HostGrid& weights, float term, int device, uint64_t* dmask,
                               float* cd, float* const cxyz[3], std::string& err) {
    const size_t n = phi.nleaf();
    if (n == 0) return 0;
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
    std::vector<int32_t> pnb, wtab;
    finest_tables(phi, weights, pnb, wtab);
    const size_t brick = n * kLeafSize * sizeof(float);
    const size_t wbytes = weights.nleaf() * kLeafSize * 3 * sizeof(float);
    DevBuf phi_d, act_d, pnb_d, w_d, wtab_d, mask_d, out_d;
    if (hipError_t e = hipMalloc(&phi_d.p, brick)) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&act_d.p, n * 8 * sizeof(uint64_t))) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&pnb_d.p, pnb.size() * sizeof(int32_t))) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&wtab_d.p, wtab.size() * sizeof(int32_t))) return hip_fail(e, "hipMalloc");
    if (wbytes)
        if (hipError_t e = hipMalloc(&w_d.p, wbytes)) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&mask_d.p, n * 8 * sizeof(uint64_t))) return hip_fail(e, "hipMalloc");
    if (hipError_t e = hipMalloc(&out_d.p, 4 * brick)) return hip_fail(e, "hipMalloc");
    const struct {
        void* dst;
        const void* src;
        size_t bytes;
    } up[] = {{phi_d.p, phi.values, brick},
              {act_d.p, phi.active.data(), n * 8 * sizeof(uint64_t)},
              {pnb_d.p, pnb.data(), pnb.size() * sizeof(int32_t)},
              {wtab_d.p, wtab.data(), wtab.size() * sizeof(int32_t)},
              {w_d.p, weights.values, wbytes}};
    for (const auto& u : up)
        if (u.bytes)
            if (hipError_t e = hipMemcpyAsync(u.dst, u.src, u.bytes, hipMemcpyHostToDevice, s))
                return hip_fail(e, "upload");
    FinestArgs a{};
    a.phi = static_cast<const float*>(phi_d.p);
    a.pact = static_cast<const uint64_t*>(act_d.p);
    a.pnb = static_cast<const int32_t*>(pnb_d.p);
    a.w = static_cast<const float*>(w_d.p);
    a.wtab = static_cast<const int32_t*>(wtab_d.p);
    a.phi_bg = phi.background[0];
    for (int c = 0; c < 3; ++c) a.w_bg[c] = weights.background[c];
    a.term = term;
    a.dmask = static_cast<uint64_t*>(mask_d.p);
    float* out = static_cast<float*>(out_d.p);
    a.cd = out;
    a.cx = out + n * kLeafSize;
    a.cy = out + 2 * n * kLeafSize;
    a.cz = out + 3 * n * kLeafSize;
    if (hipError_t e = launch_finest(int64_t(n), a, s)) return hip_fail(e, "finest kernel");
    if (hipError_t e = hipMemcpyAsync(dmask, mask_d.p, n * 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, s))
        return hip_fail(e, "download");
    float* dst[4] = {cd, cxyz[0], cxyz[1], cxyz[2]};
    for (int k = 0; k < 4; ++k)
        if (hipError_t e = hipMemcpyAsync(dst[k], out + size_t(k) * n * kLeafSize, brick, hipMemcpyDeviceToHost, s))
            return hip_fail(e, "download");
    if (hipError_t e = hipStreamSynchronize(s)) return hip_fail(e, "finest sync");
    return 0;
}

}  // namespace uaamg
