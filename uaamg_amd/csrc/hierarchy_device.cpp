// Host side of the device hierarchy build (hierarchy_device.h).
#include "hierarchy_device.h"

#include <algorithm>
#include <cmath>
#include <vector>

#include "build_timing.h"
#include "finest_device.h"

namespace uaamg {

DeviceCoefStore::~DeviceCoefStore() {
    if (blocks.empty()) return;
    int cur = -1;
    const bool had = hipGetDevice(&cur) == hipSuccess;
    if (device >= 0) (void)hipSetDevice(device);
    for (void* p : blocks)
        if (p) (void)hipFree(p);
    if (had && cur >= 0) (void)hipSetDevice(cur);
}

namespace {

constexpr size_t kAlign = 256;
size_t align_bytes(size_t b) { return (std::max<size_t>(b, 1) + kAlign - 1) & ~(kAlign - 1); }

// Lays pieces out in one allocation: a first pass with no base measures,
// the second hands out the aligned pointers.
struct Carve {
    char* base = nullptr;
    size_t used = 0;
    template <class T>
    T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
        used += align_bytes(count * sizeof(T));
        return p;
    }
};

// the device arrays of one level (kept in the store: the next coarser level's
// Galerkin sums read the bricks, DOF masks and face neighbours)
struct LevelDev {
    float* coef[4] = {nullptr, nullptr, nullptr, nullptr};
    float* invdiag = nullptr;
    uint64_t* mask = nullptr;
    int32_t* nbr = nullptr;
    int32_t* child = nullptr;
    int32_t* gather = nullptr;
    uint8_t* keep = nullptr;
};

void layout_level(Carve& c, size_t n, LevelDev& d) {
    for (float*& p : d.coef) p = c.take<float>(n * kLeafSize);
    d.invdiag = c.take<float>(n * kLeafSize);
    d.mask = c.take<uint64_t>(n * 8);
    d.nbr = c.take<int32_t>(n * 6);
    d.child = c.take<int32_t>(n * 8);
    d.gather = c.take<int32_t>(n);
    d.keep = c.take<uint8_t>(n * 4);
}

// the trimming constants of trim_level (trimDefaultNodes, simd_vdb_poisson.cpp:1756-1771)
void trim_constants(const PoissonLevelHost& l, TrimArgs& a) {
    const float term = -l.dt / (l.dx * l.dx);
    const float diagonal_term = -6.0f * term;
    const float epsilon = 1e-5f;
    a.dflt[0] = diagonal_term;
    a.eps[0] = std::fabs(diagonal_term * epsilon);
    for (int k = 1; k < 4; ++k) {
        a.dflt[k] = term;
        a.eps[k] = std::fabs(term * epsilon);
    }
    a.inv_dflt = l.default_invdiag();
}

struct Session {
    hipStream_t s = nullptr;
    std::string& err;
    std::vector<void*> temps;   // freed when the build ends
    explicit Session(std::string& e) : err(e) {}
    ~Session() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
        for (void* p : temps) (void)hipFree(p);
    }
    bool fail(hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return false;
    }
    bool up(void* dst, const void* src, size_t bytes) {
        if (!bytes) return true;
        if (hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s)) return fail(e, "upload");
        return true;
    }
    bool down(void* dst, const void* src, size_t bytes) {
        if (!bytes) return true;
        if (hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s)) return fail(e, "download");
        return true;
    }
    bool sync(const char* what) {
        if (hipError_t e = hipStreamSynchronize(s)) return fail(e, what);
        return true;
    }
};

// one level's device block in the store, its topology uploaded
bool make_level(Session& S, DeviceCoefStore& store, const PoissonLevelHost& l, LevelDev& d) {
    const size_t n = l.nleaf();
    Carve c;
    layout_level(c, n, d);
    void* p = nullptr;
    if (hipError_t e = hipMalloc(&p, c.used)) return S.fail(e, "hipMalloc");
    store.blocks.push_back(p);
    c.base = static_cast<char*>(p);
    c.used = 0;
    layout_level(c, n, d);
    if (!S.up(d.mask, l.mask.data(), n * 8 * sizeof(uint64_t))) return false;
    if (!S.up(d.nbr, l.nbr.data(), n * 6 * sizeof(int32_t))) return false;
    if (!l.child.empty() && !S.up(d.child, l.child.data(), n * 8 * sizeof(int32_t))) return false;
    return true;
}

// trims the level's bricks (gathered from src when gather is set) and fetches the keep flags
bool trim_level_device(Session& S, PoissonLevelHost& l, LevelDev& d, const float* const src[4], bool gather) {
    const size_t n = l.nleaf();
    TrimArgs a{};
    a.n = int32_t(n);
    a.gather = gather ? d.gather : nullptr;
    for (int k = 0; k < 4; ++k) {
        a.src[k] = src[k];
        a.dst[k] = d.coef[k];
    }
    a.invdiag = d.invdiag;
    a.mask = d.mask;
    trim_constants(l, a);
    a.keep = d.keep;
    if (hipError_t e = launch_level_trim(a, S.s)) return S.fail(e, "trim kernel");
    std::vector<uint8_t> keep(n * 4);
    if (!S.down(keep.data(), d.keep, n * 4)) return false;
    if (!S.sync("trim")) return false;
    const int kinds[4] = {kCoefDiag, kCoefX, kCoefY, kCoefZ};
    for (int k = 0; k < 4; ++k) l.keep[kinds[k]].assign(keep.begin() + long(k * n), keep.begin() + long((k + 1) * n));
    return true;
}

bool download_level(Session& S, PoissonLevelHost& l, const DeviceCoefStore::Level& d) {
    const size_t n = l.nleaf() * kLeafSize;
    const int kinds[4] = {kCoefDiag, kCoefX, kCoefY, kCoefZ};
    for (int k = 0; k < 4; ++k) {
        l.coef[kinds[k]].resize(n);
        if (!S.down(l.coef[kinds[k]].data(), d.coef[k], n * sizeof(float))) return false;
    }
    l.invdiag.resize(n);
    if (!S.down(l.invdiag.data(), d.invdiag, n * sizeof(float))) return false;
    if (!S.sync("download bricks")) return false;
    l.coef_on_device = false;
    return true;
}

}  // namespace

bool build_hierarchy_device(const HostGrid& phi, const HostGrid& weights, float dt, float dx,
                            int64_t max_coarsest_dof, PoissonHierarchyHost& out, int device, std::string& err) {
    StageClock clock;
    out = PoissonHierarchyHost();
    if (phi.ncomp != 1 || weights.ncomp != 3 || !(dx > 0.f)) {
        err = "bad grids";
        return false;
    }
    if (hipSetDevice(device) != hipSuccess) {
        err = "hipSetDevice failed";
        return false;
    }
    Session S(err);
    if (hipError_t e = hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking)) return S.fail(e, "hipStreamCreate");
    auto store = std::make_shared<DeviceCoefStore>();
    store->device = device;
    out.nleaf_phi = int64_t(phi.nleaf());

    PoissonLevelHost l0;
    l0.level = 0;
    l0.dt = dt;
    l0.dx = dx;
    l0.term = dt / (dx * dx);
    l0.full_origin = phi.origin;
    const size_t nphi = phi.nleaf();
    const size_t nw = weights.nleaf();

    // the per-voxel stage over every liquid-phi leaf (finest_kernel)
    std::vector<int32_t> pnb, wtab;
    finest_tables(phi, weights, pnb, wtab);
    float* phi_d = nullptr;
    float* w_d = nullptr;
    float* full[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t *act_d = nullptr, *dmask_d = nullptr;
    int32_t *pnb_d = nullptr, *wtab_d = nullptr;
    auto layout_stage = [&](Carve& c) {
        phi_d = c.take<float>(nphi * kLeafSize);
        act_d = c.take<uint64_t>(nphi * 8);
        pnb_d = c.take<int32_t>(nphi * 6);
        wtab_d = c.take<int32_t>(nphi * 4);
        w_d = c.take<float>(nw * kLeafSize * 3);
        dmask_d = c.take<uint64_t>(nphi * 8);
        for (float*& p : full) p = c.take<float>(nphi * kLeafSize);
    };
    Carve stage;
    layout_stage(stage);
    void* stage_mem = nullptr;
    if (hipError_t e = hipMalloc(&stage_mem, stage.used)) return S.fail(e, "hipMalloc");
    S.temps.push_back(stage_mem);
    stage.base = static_cast<char*>(stage_mem);
    stage.used = 0;
    layout_stage(stage);
    if (!S.up(phi_d, phi.values, nphi * kLeafSize * sizeof(float)) ||
        !S.up(act_d, phi.active.data(), nphi * 8 * sizeof(uint64_t)) ||
        !S.up(pnb_d, pnb.data(), pnb.size() * sizeof(int32_t)) ||
        !S.up(wtab_d, wtab.data(), wtab.size() * sizeof(int32_t)) ||
        !S.up(w_d, weights.values, nw * kLeafSize * 3 * sizeof(float)))
        return false;
    FinestArgs fa{};
    fa.phi = phi_d;
    fa.pact = act_d;
    fa.pnb = pnb_d;
    fa.w = w_d;
    fa.wtab = wtab_d;
    fa.phi_bg = phi.background[0];
    for (int c = 0; c < 3; ++c) fa.w_bg[c] = weights.background[c];
    fa.term = l0.term;
    fa.dmask = dmask_d;
    fa.cd = full[0];
    fa.cx = full[1];
    fa.cy = full[2];
    fa.cz = full[3];
    if (hipError_t e = launch_finest(int64_t(nphi), fa, S.s)) return S.fail(e, "finest kernel");
    std::vector<uint64_t> dmask(nphi * 8);
    if (!S.down(dmask.data(), dmask_d, dmask.size() * sizeof(uint64_t)) || !S.sync("per-voxel stage")) return false;
    clock.lap("device build: level 0 per-voxel stage");

    // compaction: the leaves holding DOFs, in tree order
    std::vector<int32_t> keep_idx;
    keep_idx.reserve(nphi);
    for (size_t i = 0; i < nphi; ++i)
        if (mask_count(&dmask[i * 8]) > 0) keep_idx.push_back(int32_t(i));
    const size_t n0 = keep_idx.size();
    l0.origin.resize(n0);
    l0.mask.resize(n0 * 8);
    for (size_t k = 0; k < n0; ++k) {
        l0.origin[k] = phi.origin[size_t(keep_idx[k])];
        std::copy(&dmask[size_t(keep_idx[k]) * 8], &dmask[size_t(keep_idx[k]) * 8] + 8, &l0.mask[k * 8]);
    }
    finalize_topology(l0);
    l0.coef_on_device = true;
    out.levels.push_back(std::move(l0));
    store->levels.emplace_back();
    LevelDev cur;
    if (!make_level(S, *store, out.levels[0], cur)) return false;
    if (!S.up(cur.gather, keep_idx.data(), n0 * sizeof(int32_t))) return false;
    if (!trim_level_device(S, out.levels[0], cur, full, true)) return false;
    // the per-voxel stage's arrays are no longer read
    (void)hipFree(stage_mem);
    S.temps.clear();
    clock.lap("device build: level 0 compaction + trimming");
    auto record = [&](size_t l, const LevelDev& d) {
        for (int k = 0; k < 4; ++k) store->levels[l].coef[k] = d.coef[k];
        store->levels[l].invdiag = d.invdiag;
    };
    record(0, cur);

    while (out.levels.back().ndof > max_coarsest_dof) {
        PoissonLevelHost next;
        build_coarse_topology(out.levels.back(), next);
        next.coef_on_device = true;
        LevelDev nd;
        if (!make_level(S, *store, next, nd)) return false;
        GalerkinArgs ga{};
        ga.n = int32_t(next.nleaf());
        ga.child = nd.child;
        ga.cmask = nd.mask;
        ga.fmask = cur.mask;
        ga.fnbr = cur.nbr;
        for (int k = 0; k < 4; ++k) {
            ga.fine[k] = cur.coef[k];
            ga.coarse[k] = nd.coef[k];
        }
        ga.term = next.term;
        if (hipError_t e = launch_galerkin(ga, S.s)) return S.fail(e, "Galerkin kernel");
        if (!trim_level_device(S, next, nd, nd.coef, false)) return false;
        out.levels.push_back(std::move(next));
        store->levels.emplace_back();
        record(out.levels.size() - 1, nd);
        cur = nd;
        clock.lap("device build: coarse level");
        if (out.levels.size() > 40) {
            err = "hierarchy does not coarsen";
            return false;
        }
    }
    // the coarsest solver forms its explicit matrix on the host
    PoissonLevelHost& last = out.levels.back();
    if (!download_level(S, last, store->levels.back())) return false;
    store->levels.back() = DeviceCoefStore::Level{};
    out.dcoef = store;
    clock.lap("device build: coarsest bricks");
    return true;
}

bool download_device_coefficients(PoissonHierarchyHost& h, std::string& err) {
    bool any = false;
    for (const auto& l : h.levels) any = any || l.coef_on_device;
    if (!any) return true;
    if (!h.dcoef) {
        err = "coefficient bricks are neither on the host nor in the device store";
        return false;
    }
    if (hipSetDevice(h.dcoef->device) != hipSuccess) {
        err = "hipSetDevice failed";
        return false;
    }
    Session S(err);
    if (hipError_t e = hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking)) return S.fail(e, "hipStreamCreate");
    for (size_t l = 0; l < h.levels.size(); ++l) {
        if (!h.levels[l].coef_on_device) continue;
        if (l >= h.dcoef->levels.size() || !h.dcoef->levels[l].coef[0]) {
            err = "level bricks missing from the device store";
            return false;
        }
        if (!download_level(S, h.levels[l], h.dcoef->levels[l])) return false;
    }
    return true;
}

}  // namespace uaamg
