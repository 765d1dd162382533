// C ABI of the variational viscosity solve (include/uaamg_capi.h, viscosity section).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <memory>
#include <vector>

#include "capi_common.h"
#include "scene.h"
#include "visc_batch.h"
#include "viscosity_solver.h"

using namespace uaamg;
using uaamg::capi::fail;
using uaamg::capi::to_host_grid;

struct uaamg_viscosity {
    ViscDevice dev;
    float* dofbuf[3] = {nullptr, nullptr, nullptr};
    int64_t ndof = 0;
    ~uaamg_viscosity() {
        for (auto p : dofbuf)
            if (p) (void)hipFree(p);
    }
};

struct uaamg_visc_scene {
    ViscScene s;
};

namespace {
int dev_status(int rc, const ViscDevice& d) {
    if (rc == -4) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    if (rc == -1) return fail(UAAMG_EINVAL, d.error().empty() ? "bad argument" : d.error());
    if (rc == -2) return fail(UAAMG_ENOMEM, d.error().empty() ? "device allocation failed" : d.error());
    if (rc < 0) return fail(UAAMG_EHIP, d.error().empty() ? "device error" : d.error());
    return rc;
}
int64_t lvl_ndof(const uaamg_viscosity* h, int level) { return h->dev.host().levels[size_t(level)].ndof; }
bool bad_level(const uaamg_viscosity* h, int level) { return !h || level < 0 || level >= h->dev.nlevels(); }

int copy_in(uaamg_viscosity* h, int slot, const float* src, int64_t n) {
    if (!src) return 0;
    hipError_t e = hipMemcpyAsync(h->dofbuf[slot], src, size_t(n) * 4, hipMemcpyHostToDevice, h->dev.stream());
    return e == hipSuccess ? 0 : fail(UAAMG_EHIP, hipGetErrorString(e));
}
int copy_out(uaamg_viscosity* h, int slot, float* dst, int64_t n) {
    hipError_t e = hipMemcpyAsync(dst, h->dofbuf[slot], size_t(n) * 4, hipMemcpyDeviceToHost, h->dev.stream());
    if (e == hipSuccess) e = hipStreamSynchronize(h->dev.stream());
    return e == hipSuccess ? 0 : fail(UAAMG_EHIP, hipGetErrorString(e));
}
}  // namespace

extern "C" {

int uaamg_viscosity_options_default(uaamg_viscosity_options* o) {
    if (!o) return fail(UAAMG_EINVAL, "null options");
    o->max_iterations = 100;
    o->device = 0;
    o->max_coarsest_dof = 4000;
    o->flags = 0;
    o->precision = UAAMG_PRECISION_FP32;
    o->stream = nullptr;
    return 0;
}

}  // extern "C"

namespace uaamg {
namespace capi {
int prepare_viscosity(const uaamg_grid* viscosity, const uaamg_grid* liquid_sdf, const uaamg_grid* solid_sdf,
                      float dt, float rho, float dx, const uaamg_viscosity_options* opts, ViscHierarchyHost& hh,
                      ViscOptions& vo) {
    if (!(dx > 0.f) || !(dt > 0.f) || !(rho > 0.f)) return fail(UAAMG_EINVAL, "dt, rho and dx must be positive");
    uaamg_viscosity_options o;
    uaamg_viscosity_options_default(&o);
    if (opts) o = *opts;
    if (o.max_coarsest_dof < 1 || o.max_coarsest_dof > 4096)
        return fail(UAAMG_EINVAL, "max_coarsest_dof must lie in [1, 4096]");
    if (o.precision != UAAMG_PRECISION_FP32 && o.precision != UAAMG_PRECISION_BF16_CYCLE)
        return fail(UAAMG_EINVAL, "unknown precision");
    if (o.device >= 0) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= o.device)
            return fail(UAAMG_ENODEV, "no HIP device with that ordinal");
    }
    HostGrid mu, liq, sol;
    if (!to_host_grid(viscosity, 1, mu, "viscosity")) return UAAMG_EINVAL;
    if (!to_host_grid(liquid_sdf, 1, liq, "liquid_sdf")) return UAAMG_EINVAL;
    if (!to_host_grid(solid_sdf, 1, sol, "solid_sdf")) return UAAMG_EINVAL;
    // the level-0 volume stage runs on the handle's device unless flags bit 8 keeps it on the host
    const int volume_device = (o.device >= 0 && !(o.flags & 256)) ? o.device : -1;
    std::string err;
    if (!build_visc_hierarchy(liq, sol, mu, dt, rho, dx, o.max_coarsest_dof, !(o.flags & 8), hh, volume_device, &err))
        return fail(UAAMG_EINVAL, "hierarchy construction failed: " + err);
    if (hh.levels.empty() || hh.levels[0].ndof == 0) return fail(UAAMG_EINVAL, "no degrees of freedom");
    vo = ViscOptions{};
    vo.max_iterations = o.max_iterations;
    vo.device = o.device;
    vo.max_coarsest_dof = o.max_coarsest_dof;
    vo.flags = o.flags;
    vo.precision = o.precision;
    vo.stream = static_cast<hipStream_t>(o.stream);
    return 0;
}
}  // namespace capi
}  // namespace uaamg

extern "C" {

int uaamg_viscosity_create(const uaamg_grid* viscosity, const uaamg_grid* liquid_sdf, const uaamg_grid* solid_sdf,
                           float dt, float rho, float dx, const uaamg_viscosity_options* opts,
                           uaamg_viscosity** out) {
    UAAMG_GUARD_BEGIN
    if (!out) return fail(UAAMG_EINVAL, "null output handle");
    *out = nullptr;
    ViscHierarchyHost hh;
    ViscOptions vo;
    if (int rc = capi::prepare_viscosity(viscosity, liquid_sdf, solid_sdf, dt, rho, dx, opts, hh, vo)) return rc;
    std::unique_ptr<uaamg_viscosity> h(new uaamg_viscosity());
    h->ndof = hh.levels[0].ndof;
    if (vo.device < 0) {
        h->dev.init_host_only(std::move(hh), vo);
    } else {
        int rc = h->dev.init(std::move(hh), vo);
        if (rc) return dev_status(rc, h->dev);
        for (auto& p : h->dofbuf)
            if (hipMalloc(reinterpret_cast<void**>(&p), size_t(std::max<int64_t>(h->ndof, 1)) * 4) != hipSuccess)
                return fail(UAAMG_ENOMEM, "device allocation failed");
    }
    *out = h.release();
    return 0;
    UAAMG_GUARD_END
}

void uaamg_viscosity_destroy(uaamg_viscosity* h) { delete h; }

int uaamg_viscosity_kernel_map(const uaamg_viscosity* h, int level, int32_t* small, float* ms_small, float* ms_pass) {
    if (!h || level < 0 || level >= h->dev.nlevels()) return fail(UAAMG_EINVAL, "bad level");
    const auto& m = h->dev.small_levels();
    if (small) *small = level < int(m.size()) ? m[size_t(level)] : 0;
    if (ms_small) *ms_small = h->dev.tune_ms(level, 0);
    if (ms_pass) *ms_pass = h->dev.tune_ms(level, 1);
    return 0;
}

int uaamg_viscosity_set_options(uaamg_viscosity* h, const uaamg_viscosity_options* o) {
    if (!h || !o) return fail(UAAMG_EINVAL, "null argument");
    if (o->max_iterations < 0) return fail(UAAMG_EINVAL, "max_iterations must be >= 0");
    h->dev.options().max_iterations = o->max_iterations;
    return 0;
}

int64_t uaamg_viscosity_ndof(const uaamg_viscosity* h) { return h ? h->ndof : -1; }
int uaamg_viscosity_nlevels(const uaamg_viscosity* h) { return h ? h->dev.nlevels() : -1; }

int uaamg_viscosity_level_info(const uaamg_viscosity* h, int level, int64_t* ndof, int64_t* nchan3,
                               int64_t* nleaf_dof) {
    if (bad_level(h, level)) return fail(UAAMG_EINVAL, "bad handle or level");
    const ViscLevelHost& L = h->dev.host().levels[size_t(level)];
    if (ndof) *ndof = L.ndof;
    if (nchan3)
        for (int c = 0; c < 3; ++c) nchan3[c] = L.nchan[c];
    if (nleaf_dof) *nleaf_dof = int64_t(L.nleaf_dof);
    return 0;
}

int uaamg_viscosity_level_topology(const uaamg_viscosity* h, int level, int32_t* origins, uint64_t* masks,
                                   int64_t* dofbase) {
    if (bad_level(h, level)) return fail(UAAMG_EINVAL, "bad handle or level");
    const ViscLevelHost& L = h->dev.host().levels[size_t(level)];
    for (size_t i = 0; i < L.nleaf_dof; ++i) {
        if (origins) {
            origins[3 * i] = L.origin[i].x;
            origins[3 * i + 1] = L.origin[i].y;
            origins[3 * i + 2] = L.origin[i].z;
        }
    }
    if (masks) std::memcpy(masks, L.mask.data(), L.mask.size() * 8);
    if (dofbase) std::memcpy(dofbase, L.dofbase.data(), L.dofbase.size() * 8);
    return 0;
}

int uaamg_viscosity_level_coefficients(uaamg_viscosity* h, int level, int grid, float* bricks, uint8_t* kept) {
    UAAMG_GUARD_BEGIN
    if (bad_level(h, level) || grid < 0 || grid >= kVNumGrids) return fail(UAAMG_EINVAL, "bad handle, level or grid");
    ViscLevelHost& L = h->dev.host_mut().levels[size_t(level)];
    if (grid >= kVSpai0) build_visc_spai0(L);
    for (size_t i = 0; i < L.nleaf_dof; ++i) {
        const int32_t s = L.slot[i * kVSlots + grid];
        if (kept) kept[i] = s >= 0 ? 1 : 0;
        if (bricks)
            for (int off = 0; off < 512; ++off) bricks[i * 512 + off] = L.coef(grid, i, off);
    }
    return 0;
    UAAMG_GUARD_END
}

int uaamg_viscosity_level_matrix(const uaamg_viscosity* h, int level, int64_t* nnz, int64_t* ptr, int32_t* col,
                                 float* val) {
    UAAMG_GUARD_BEGIN
    if (bad_level(h, level)) return fail(UAAMG_EINVAL, "bad handle or level");
    std::vector<int64_t> p;
    std::vector<int32_t> c;
    std::vector<float> v;
    visc_explicit_matrix(h->dev.host().levels[size_t(level)], p, c, v);
    if (nnz) *nnz = int64_t(c.size());
    if (ptr) std::memcpy(ptr, p.data(), p.size() * 8);
    if (col) std::memcpy(col, c.data(), c.size() * 4);
    if (val) std::memcpy(val, v.data(), v.size() * 4);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_viscosity_build_rhs(const uaamg_viscosity* h, const uaamg_grid* velocity, const uaamg_grid* solid_velocity,
                              float* rhs) {
    UAAMG_GUARD_BEGIN
    if (!h || !rhs) return fail(UAAMG_EINVAL, "null argument");
    HostGrid u, us;
    if (!to_host_grid(velocity, 3, u, "velocity")) return UAAMG_EINVAL;
    if (solid_velocity && !to_host_grid(solid_velocity, 3, us, "solid_velocity")) return UAAMG_EINVAL;
    build_visc_rhs(h->dev.host(), u, solid_velocity ? &us : nullptr, rhs);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_viscosity_build_rhs_packed(const uaamg_viscosity* h, const uaamg_grid* v0, const uaamg_grid* v1,
                                     const uaamg_grid* v2, const uaamg_grid* solid_velocity, float* rhs) {
    UAAMG_GUARD_BEGIN
    if (!h || !v0 || !v1 || !v2 || !rhs) return fail(UAAMG_EINVAL, "null argument");
    HostGrid u[3], us;
    const uaamg_grid* in[3] = {v0, v1, v2};
    const char* names[3] = {"velocity channel 0", "velocity channel 1", "velocity channel 2"};
    for (int c = 0; c < 3; ++c)
        if (!to_host_grid(in[c], 1, u[c], names[c])) return UAAMG_EINVAL;
    if (solid_velocity && !to_host_grid(solid_velocity, 3, us, "solid_velocity")) return UAAMG_EINVAL;
    const HostGrid* ch[3] = {&u[0], &u[1], &u[2]};
    build_visc_rhs(h->dev.host(), ch, solid_velocity ? &us : nullptr, rhs);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_viscosity_grid_to_vector(const uaamg_viscosity* h, const uaamg_grid* v0, const uaamg_grid* v1,
                                   const uaamg_grid* v2, float* vec) {
    UAAMG_GUARD_BEGIN
    if (!h || !v0 || !v1 || !v2 || !vec) return fail(UAAMG_EINVAL, "null argument");
    HostGrid g[3];
    const uaamg_grid* in[3] = {v0, v1, v2};
    for (int c = 0; c < 3; ++c)
        if (!to_host_grid(in[c], 1, g[c], "velocity channel")) return UAAMG_EINVAL;
    const ViscLevelHost& L = h->dev.host().levels[0];
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < L.nleaf_dof; ++i) {
            int64_t d = L.dofbase[i * 3 + size_t(c)];
            if (d < 0) continue;
            const int32_t gi = g[c].lookup.find(L.origin[i]);
            const uint64_t* m = &L.mask[(i * 3 + size_t(c)) * 8];
            for (int off = 0; off < kLeafSize; ++off)
                if (mask_test(m, off)) vec[d++] = gi < 0 ? 0.f : g[c].leaf_values(size_t(gi))[off];
        }
    return 0;
    UAAMG_GUARD_END
}

int uaamg_viscosity_vector_to_grid(const uaamg_viscosity* h, const float* vec, float* bricks) {
    UAAMG_GUARD_BEGIN
    if (!h || !vec || !bricks) return fail(UAAMG_EINVAL, "null argument");
    const ViscLevelHost& L = h->dev.host().levels[0];
    for (size_t i = 0; i < L.nleaf_dof; ++i)
        for (int c = 0; c < 3; ++c) {
            int64_t d = L.dofbase[i * 3 + size_t(c)];
            const uint64_t* m = &L.mask[(i * 3 + size_t(c)) * 8];
            float* out = bricks + (i * 3 + size_t(c)) * kLeafSize;
            for (int off = 0; off < kLeafSize; ++off) out[off] = (d >= 0 && mask_test(m, off)) ? vec[d++] : 0.f;
        }
    return 0;
    UAAMG_GUARD_END
}

int uaamg_viscosity_solve_pcg_device(uaamg_viscosity* h, float tolerance, float* x_dev, const float* b_dev,
                                     int32_t* iterations) {
    if (!h || !x_dev || !b_dev) return fail(UAAMG_EINVAL, "null argument");
    h->dev.options().tolerance = tolerance;
    int it = 0;
    const int rc = h->dev.solve_pcg_dev(x_dev, b_dev, &it);
    if (iterations) *iterations = it;
    return rc < 0 ? dev_status(rc, h->dev) : 0;
}

int uaamg_viscosity_solve_batch_device(uaamg_viscosity* const* hs, int32_t n, float tolerance, float* const* x_dev,
                                       const float* const* b_dev, int32_t* iterations, int32_t concurrency) {
    UAAMG_GUARD_BEGIN
    if (n < 0 || (n > 0 && (!hs || !x_dev || !b_dev))) return fail(UAAMG_EINVAL, "null argument");
    for (int32_t i = 0; i < n; ++i) {
        if (!hs[i] || !x_dev[i] || !b_dev[i]) return fail(UAAMG_EINVAL, "null handle or vector in the batch");
        if (!hs[i]->dev.on_device()) return fail(UAAMG_ENODEV, "a batch handle was created without a device");
        for (int32_t j = 0; j < i; ++j)
            if (hs[j] == hs[i]) return fail(UAAMG_EINVAL, "a handle appears twice in the batch");
    }
    const int nt = std::max(1, std::min<int>(std::min<int>(concurrency, 16), n));
    std::atomic<int32_t> next{0};
    std::vector<int> rcs(size_t(std::max(n, 1)), 0);
    std::vector<std::string> errs(size_t(std::max(n, 1)));
    auto worker = [&]() {
        for (;;) {
            const int32_t i = next.fetch_add(1);
            if (i >= n) return;
            ViscDevice& d = hs[i]->dev;
            d.options().tolerance = tolerance;
            int it = 0;
            int rc = hipSetDevice(d.options().device) == hipSuccess ? d.solve_pcg_dev(x_dev[i], b_dev[i], &it) : -3;
            rcs[size_t(i)] = rc;
            if (iterations) iterations[i] = it;
            if (rc < 0) errs[size_t(i)] = d.error().empty() ? "device error" : d.error();
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    for (int32_t i = 0; i < n; ++i)
        if (rcs[size_t(i)] < 0) return fail(UAAMG_EHIP, "batch problem " + std::to_string(i) + ": " + errs[size_t(i)]);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_viscosity_solve_pcg(uaamg_viscosity* h, float tolerance, float* x, const float* b, int32_t* iterations) {
    if (!h || !x || !b) return fail(UAAMG_EINVAL, "null argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    if (int rc = copy_in(h, 0, x, h->ndof)) return rc;
    if (int rc = copy_in(h, 1, b, h->ndof)) return rc;
    const int rc = uaamg_viscosity_solve_pcg_device(h, tolerance, h->dofbuf[0], h->dofbuf[1], iterations);
    if (rc < 0) return rc;
    if (int e = copy_out(h, 0, x, h->ndof)) return e;
    return rc;
}

int uaamg_viscosity_apply(uaamg_viscosity* h, int level, int mode, const float* x, const float* b, float* y) {
    if (bad_level(h, level) || !x || !y || mode < 0 || mode > 7) return fail(UAAMG_EINVAL, "bad argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    const int64_t n = lvl_ndof(h, level);
    if (int rc = copy_in(h, 0, x, n)) return rc;
    if (int rc = copy_in(h, 1, b, n)) return rc;
    const int rc = h->dev.apply_dev(level, mode, h->dofbuf[0], b ? h->dofbuf[1] : nullptr, h->dofbuf[2]);
    if (rc) return dev_status(rc, h->dev);
    return copy_out(h, 2, y, n);
}

int uaamg_viscosity_restrict(uaamg_viscosity* h, int fine_level, const float* fine, float* coarse) {
    if (bad_level(h, fine_level) || fine_level + 1 >= h->dev.nlevels() || !fine || !coarse)
        return fail(UAAMG_EINVAL, "bad argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    if (int rc = copy_in(h, 0, fine, lvl_ndof(h, fine_level))) return rc;
    const int rc = h->dev.restrict_dev(fine_level, h->dofbuf[0], h->dofbuf[1]);
    if (rc) return dev_status(rc, h->dev);
    return copy_out(h, 1, coarse, lvl_ndof(h, fine_level + 1));
}

int uaamg_viscosity_prolong(uaamg_viscosity* h, int fine_level, const float* coarse, float* fine) {
    if (bad_level(h, fine_level) || fine_level + 1 >= h->dev.nlevels() || !fine || !coarse)
        return fail(UAAMG_EINVAL, "bad argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    if (int rc = copy_in(h, 0, coarse, lvl_ndof(h, fine_level + 1))) return rc;
    if (int rc = copy_in(h, 1, fine, lvl_ndof(h, fine_level))) return rc;
    const int rc = h->dev.prolong_dev(fine_level, h->dofbuf[0], h->dofbuf[1]);
    if (rc) return dev_status(rc, h->dev);
    return copy_out(h, 1, fine, lvl_ndof(h, fine_level));
}

int uaamg_viscosity_coarsest_solve(uaamg_viscosity* h, const float* b, float* x) {
    if (!h || !b || !x) return fail(UAAMG_EINVAL, "null argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    const int64_t n = lvl_ndof(h, h->dev.nlevels() - 1);
    if (int rc = copy_in(h, 0, b, n)) return rc;
    const int rc = h->dev.coarsest_dev(h->dofbuf[0], h->dofbuf[1]);
    if (rc) return dev_status(rc, h->dev);
    return copy_out(h, 1, x, n);
}

int uaamg_viscosity_precondition(uaamg_viscosity* h, const float* r, float* z) {
    if (!h || !r || !z) return fail(UAAMG_EINVAL, "null argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    if (int rc = copy_in(h, 0, r, h->ndof)) return rc;
    const int rc = h->dev.precondition_dev(h->dofbuf[0], h->dofbuf[1]);
    if (rc) return dev_status(rc, h->dev);
    return copy_out(h, 1, z, h->ndof);
}

int uaamg_viscosity_algorithmic_bytes(const uaamg_viscosity* h, int level, int mode, int channel, double* bytes) {
    if (bad_level(h, level) || !bytes || mode < 0 || mode > 5 || channel < 0 || channel > 2)
        return fail(UAAMG_EINVAL, "bad argument");
    *bytes = h->dev.algorithmic_bytes(level, mode, channel);
    return 0;
}

int uaamg_viscosity_time_apply(uaamg_viscosity* h, int level, int mode, int channel, int reps, float* ms) {
    if (!h || !ms) return fail(UAAMG_EINVAL, "null argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    const int rc = h->dev.time_apply(level, mode, channel, reps, ms);
    return rc ? dev_status(rc, h->dev) : 0;
}

int uaamg_visc_scene_create(int config, int32_t n, uint64_t seed, float mu, float dt, float rho,
                            uaamg_visc_scene** out) {
    UAAMG_GUARD_BEGIN
    if (!out) return fail(UAAMG_EINVAL, "null output");
    SceneSpec spec;
    if (!scene_preset(config, n, seed, spec)) return fail(UAAMG_EINVAL, "unknown scene config");
    std::unique_ptr<uaamg_visc_scene> s(new uaamg_visc_scene());
    build_visc_scene(spec, mu, dt, rho, s->s);
    *out = s.release();
    return 0;
    UAAMG_GUARD_END
}

int uaamg_visc_scene_grids(const uaamg_visc_scene* sc, uaamg_grid* liquid, uaamg_grid* solid, uaamg_grid* viscosity,
                           uaamg_grid* velocity, float* dx, float* dt, float* rho) {
    if (!sc) return fail(UAAMG_EINVAL, "null scene");
    const ViscScene& s = sc->s;
    if (liquid) {
        liquid->nleaf = s.liquid.nleaf();
        liquid->origins = s.liquid.origins.data();
        liquid->active = s.liquid.active.data();
        liquid->values = s.liquid.phi.data();
        liquid->ncomp = 1;
        liquid->background[0] = liquid->background[1] = liquid->background[2] = s.liquid.phi_background;
    }
    if (solid) {
        solid->nleaf = int64_t(s.solid_origins.size() / 3);
        solid->origins = s.solid_origins.data();
        solid->active = s.solid_active.data();
        solid->values = s.solid_phi.data();
        solid->ncomp = 1;
        solid->background[0] = solid->background[1] = solid->background[2] = s.solid_background;
    }
    if (viscosity) {
        viscosity->nleaf = 0;
        viscosity->origins = nullptr;
        viscosity->active = nullptr;
        viscosity->values = nullptr;
        viscosity->ncomp = 1;
        viscosity->background[0] = viscosity->background[1] = viscosity->background[2] = s.mu;
    }
    if (velocity) {
        velocity->nleaf = int64_t(s.vel_origins.size() / 3);
        velocity->origins = s.vel_origins.data();
        velocity->active = nullptr;
        velocity->values = s.vel.data();
        velocity->ncomp = 3;
        velocity->background[0] = velocity->background[1] = velocity->background[2] = 0.f;
    }
    if (dx) *dx = s.liquid.dx;
    if (dt) *dt = s.dt;
    if (rho) *rho = s.rho;
    return 0;
}

void uaamg_visc_scene_destroy(uaamg_visc_scene* s) { delete s; }

}  // extern "C"

struct uaamg_viscosity_batch {
    ViscBatch b;
};

extern "C" {

int uaamg_viscosity_batch_create(int32_t n, const uaamg_grid* const* viscosity, const uaamg_grid* const* liquid_sdf,
                                 const uaamg_grid* const* solid_sdf, const float* dt, const float* rho,
                                 const float* dx, const uaamg_viscosity_options* opts,
                                 uaamg_viscosity_batch** out) {
    UAAMG_GUARD_BEGIN
    if (!out) return fail(UAAMG_EINVAL, "null output handle");
    *out = nullptr;
    if (n <= 0 || !viscosity || !liquid_sdf || !solid_sdf || !dt || !rho || !dx)
        return fail(UAAMG_EINVAL, "empty or null batch");
    std::vector<ViscHierarchyHost> probs(static_cast<size_t>(n));
    ViscOptions vo;
    for (int32_t i = 0; i < n; ++i) {
        ViscOptions oi;
        if (int rc = capi::prepare_viscosity(viscosity[i], liquid_sdf[i], solid_sdf[i], dt[i], rho[i], dx[i], opts,
                                             probs[size_t(i)], oi))
            return rc;
        if (i == 0) vo = oi;
    }
    if (vo.device < 0) return fail(UAAMG_ENODEV, "the merged batch needs a device");
    std::unique_ptr<uaamg_viscosity_batch> h(new uaamg_viscosity_batch());
    int rc = h->b.init(std::move(probs), vo);
    if (rc < 0) {
        const int code = rc == -1 ? UAAMG_EINVAL : rc == -2 ? UAAMG_ENOMEM : rc == -4 ? UAAMG_ENODEV : UAAMG_EHIP;
        return fail(code, h->b.error().empty() ? "batch initialisation failed" : h->b.error());
    }
    *out = h.release();
    return 0;
    UAAMG_GUARD_END
}

void uaamg_viscosity_batch_destroy(uaamg_viscosity_batch* h) { delete h; }

int uaamg_viscosity_batch_info(const uaamg_viscosity_batch* h, int32_t* nproblems, int32_t* nlevels,
                               int64_t* leaves0, int64_t* device_bytes) {
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    if (nproblems) *nproblems = h->b.nproblems();
    if (nlevels) *nlevels = h->b.nlevels();
    if (leaves0) *leaves0 = h->b.level_leaves(0);
    if (device_bytes) *device_bytes = int64_t(h->b.device_bytes());
    return 0;
}

int64_t uaamg_viscosity_batch_ndof(const uaamg_viscosity_batch* h, int32_t i) { return h ? h->b.ndof(i) : -1; }

int uaamg_viscosity_batch_solve_device(uaamg_viscosity_batch* h, float tolerance, float* const* x_dev,
                                       const float* const* b_dev, int32_t* iterations) {
    UAAMG_GUARD_BEGIN
    if (!h || !x_dev || !b_dev) return fail(UAAMG_EINVAL, "null argument");
    const int n = h->b.nproblems();
    std::vector<int> it(static_cast<size_t>(n));
    int rc = h->b.solve(tolerance, x_dev, b_dev, it.data());
    if (rc < 0) {
        const int code = rc == -1 ? UAAMG_EINVAL : rc == -4 ? UAAMG_ENODEV : UAAMG_EHIP;
        return fail(code, h->b.error().empty() ? "batch solve failed" : h->b.error());
    }
    if (iterations)
        for (int i = 0; i < n; ++i) iterations[i] = it[size_t(i)];
    return 0;
    UAAMG_GUARD_END
}

float uaamg_volume_fraction8(const float* p) {
    return p ? visc_volume_fraction(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]) : 0.f;
}

}  // extern "C"
