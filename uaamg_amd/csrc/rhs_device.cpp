// Host side of the device right-hand-side build (rhs_device.h).
#include "rhs_device.h"

#include <mutex>
#include <vector>

#include "leaf_topology.h"
#include "poisson_solver.h"

namespace uaamg {

// Scratch of the allocation-free entry point: the leaf tables of one call
// ([nleaf0][kRhsGrids][4] int32, 80 bytes per level-0 DOF leaf). Every GPU the
// library runs on holds its own copy until the process ends. In host code the
// name is only a stand-in for the symbol; kernels get the per-device address
// that hipGetSymbolAddress returns.
static __device__ __align__(256) int32_t scratch[size_t(kRhsNoAllocMaxLeaves) * kRhsGrids * 4];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on `device`, looked up once per device with
// that device current (the caller has made it current).
int scratch_address(int device, int32_t** out, std::string& err) {
    if (device < 0 || device >= kMaxScratchDevices) {
        err = "device ordinal out of range for the scratch table";
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return -3;
        }
        g_scratch_ptr[device] = static_cast<int32_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return 0;
}

// device memory freed at scope exit (the call waits for its kernel first)
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// tab_scratch: device memory for the leaf tables (at least kRhsNoAllocMaxLeaves
// leaves' worth), or null to allocate them for the call
int build_rhs_impl(PoissonDevice& d, const RhsGridRef* grids, float* rhs_dof, int32_t* tab_scratch,
                   std::string& err) {
    if (!d.on_device()) {
        err = "handle was created without a device (device < 0)";
        return -4;
    }
    const PoissonLevelHost& l0 = d.host().levels[0];
    const size_t n0 = l0.nleaf();
    hipStream_t s = d.stream();
    if (hipSetDevice(d.device()) != hipSuccess) {
        err = "hipSetDevice failed";
        return -3;
    }
    if (d.follow_caller()) {
        err = d.error();
        return -3;
    }
    auto hip_fail = [&](hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        (void)hipStreamSynchronize(s);   // no copy still reads a buffer that goes out of scope
        return -3;
    };
    RhsArgs a{};
    a.invdx = 1.0f / l0.dx;
    DevBuf vals[kRhsGrids], tab;
    std::vector<int32_t> table(n0 * kRhsGrids * 4, -1);
    for (int g = 0; g < kRhsGrids; ++g) {
        const RhsGridRef& r = grids[g];
        std::vector<Coord> org(size_t(r.nleaf));
        for (int64_t i = 0; i < r.nleaf; ++i) {
            org[size_t(i)] = {r.origins[3 * i], r.origins[3 * i + 1], r.origins[3 * i + 2]};
            if ((org[size_t(i)].x | org[size_t(i)].y | org[size_t(i)].z) & 7) {
                err = "grid leaf origin not a multiple of 8";
                return -1;
            }
        }
        LeafLookup lookup;
        lookup.build(org);
        for (size_t i = 0; i < org.size(); ++i)
            if (lookup.find(org[i]) != int32_t(i)) {
                err = "grid has duplicate leaf origins";
                return -1;
            }
        for (size_t i = 0; i < n0; ++i) {
            const Coord o = l0.origin[i];
            const Coord at[4] = {o, {o.x + 8, o.y, o.z}, {o.x, o.y + 8, o.z}, {o.x, o.y, o.z + 8}};
            for (int k = 0; k < 4; ++k) table[(i * kRhsGrids + size_t(g)) * 4 + size_t(k)] = lookup.find(at[k]);
        }
        a.g[g].ncomp = r.ncomp;
        for (int c = 0; c < 3; ++c) a.g[g].bg[c] = r.bg[c];
        const size_t bytes = size_t(r.nleaf) * kLeafSize * size_t(r.ncomp) * sizeof(float);
        if (r.nleaf == 0) {
            a.g[g].values = nullptr;   // every lookup misses
        } else if (r.values_on_device) {
            a.g[g].values = r.values;
        } else {
            if (hipError_t e = hipMalloc(&vals[g].p, bytes)) return hip_fail(e, "hipMalloc");
            if (hipError_t e = hipMemcpyAsync(vals[g].p, r.values, bytes, hipMemcpyHostToDevice, s))
                return hip_fail(e, "upload grid values");
            a.g[g].values = static_cast<const float*>(vals[g].p);
        }
    }
    if (n0 == 0) return 0;
    void* tab_dev = tab_scratch;
    if (!tab_dev) {
        if (hipError_t e = hipMalloc(&tab.p, table.size() * sizeof(int32_t))) return hip_fail(e, "hipMalloc");
        tab_dev = tab.p;
    }
    if (hipError_t e = hipMemcpyAsync(tab_dev, table.data(), table.size() * sizeof(int32_t), hipMemcpyHostToDevice, s))
        return hip_fail(e, "upload leaf tables");
    a.tab = static_cast<const int32_t*>(tab_dev);
    if (hipError_t e = launch_poisson_rhs(d.level_view(0), a, rhs_dof, s)) return hip_fail(e, "rhs kernel");
    if (hipError_t e = hipStreamSynchronize(s)) return hip_fail(e, "rhs sync");
    return 0;
}

}  // namespace

int build_poisson_rhs_device(PoissonDevice& d, const RhsGridRef* grids, float* rhs_dof, std::string& err) {
    return build_rhs_impl(d, grids, rhs_dof, nullptr, err);
}

int build_poisson_rhs_device_noalloc(PoissonDevice& d, const RhsGridRef* grids, float* rhs_dof, std::string& err) {
    if (!d.on_device()) {
        err = "handle was created without a device (device < 0)";
        return -4;
    }
    if (hipSetDevice(d.device()) != hipSuccess) {
        err = "hipSetDevice failed";
        return -3;
    }
    for (int g = 0; g < kRhsGrids; ++g) {
        if (grids[g].nleaf == 0) continue;
        // the values must be device memory: a host pointer handed to the kernel would fault
        hipPointerAttribute_t at{};
        const hipError_t e = hipPointerGetAttributes(&at, grids[g].values);
        if (!grids[g].values_on_device || e != hipSuccess ||
            (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)) {
            (void)hipGetLastError();
            err = "the allocation-free build reads every grid's values from device memory";
            return -1;
        }
    }
    const int64_t n0 = int64_t(d.host().levels[0].nleaf());
    if (n0 > kRhsNoAllocMaxLeaves) {
        err = "level 0 has " + std::to_string(n0) + " DOF leaves; the allocation-free build takes at most " +
              std::to_string(kRhsNoAllocMaxLeaves);
        return -1;
    }
    int32_t* tab = nullptr;
    if (int rc = scratch_address(d.device(), &tab, err)) return rc;
    return build_rhs_impl(d, grids, rhs_dof, tab, err);
}

}  // namespace uaamg
