// C ABI over the device solver (include/uaamg_capi.h).
#include "../../include/uaamg_capi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>
#include <memory>
#include <new>
#include <string>

#include "batch_solver.h"
#include "poisson_host.h"
#include "poisson_solver.h"
#include "rhs_device.h"
#include "face_update.h"
#include "face_weights.h"
#include "hierarchy_device.h"
#include "build_timing.h"
#include "capi_common.h"
#include "scene.h"

using namespace uaamg;

struct uaamg_poisson {
    PoissonDevice dev;
    float* dofbuf[3] = {nullptr, nullptr, nullptr};   // DOF-ordered device vectors for the host API
    int64_t ndof = 0;
    ~uaamg_poisson() {
        for (auto p : dofbuf)
            if (p) (void)hipFree(p);
    }
};

struct uaamg_scene {
    Scene s;
};

namespace uaamg {
namespace capi {
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

bool to_host_grid(const uaamg_grid* g, int ncomp, HostGrid& out, const char* what) {
    if (!g) {
        g_err = std::string(what) + ": null grid";
        return false;
    }
    if (g->ncomp != ncomp) {
        g_err = std::string(what) + ": wrong component count";
        return false;
    }
    // the grids built here live only for the entry point's call, so tree-ordered
    // caller values are read in place
    if (!make_host_grid(g->nleaf, g->origins, g->active, g->values, g->ncomp, g->background, out, true)) {
        g_err = std::string(what) + ": malformed leaves (origin not a multiple of 8, duplicate, or null arrays)";
        return false;
    }
    return true;
}
}  // namespace capi
}  // namespace uaamg

using namespace uaamg::capi;

namespace {
int hip_status(int rc, const PoissonDevice& d) {
    if (rc == -4) g_err = "handle was created without a device (device < 0)";
    else if (rc < 0) g_err = d.error().empty() ? "device error" : d.error();
    return rc;
}

int copy_in(uaamg_poisson* h, int slot, const float* src, int64_t n) {
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    if (!src) return 0;
    hipError_t e = hipMemcpyAsync(h->dofbuf[slot], src, size_t(n) * sizeof(float), hipMemcpyHostToDevice,
                                  h->dev.stream());
    return e == hipSuccess ? 0 : fail(UAAMG_EHIP, hipGetErrorString(e));
}
int copy_out(uaamg_poisson* h, int slot, float* dst, int64_t n) {
    hipError_t e = hipMemcpyAsync(dst, h->dofbuf[slot], size_t(n) * sizeof(float), hipMemcpyDeviceToHost,
                                  h->dev.stream());
    if (e == hipSuccess) e = hipStreamSynchronize(h->dev.stream());
    return e == hipSuccess ? 0 : fail(UAAMG_EHIP, hipGetErrorString(e));
}
int64_t level_ndof(const uaamg_poisson* h, int level) { return h->dev.host().levels[level].ndof; }
}  // namespace

extern "C" {

const char* uaamg_last_error(void) { return g_err.c_str(); }

const char* uaamg_version(void) { return "uaamg_amd 0.1 (gfx950)"; }

int uaamg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int uaamg_options_default(uaamg_options* o) {
    if (!o) return fail(UAAMG_EINVAL, "null options");
    std::memset(o, 0, sizeof(*o));
    o->relative_tolerance = 1e-7f;
    o->max_iterations = 100;
    o->smoother = UAAMG_SMOOTHER_SCHEDULED_RELAXED_JACOBI;
    o->precision = 0;
    o->device = 0;
    o->max_coarsest_dof = 4000;
    o->stream = nullptr;
    return 0;
}

}  // extern "C"

namespace uaamg {
namespace capi {
int prepare_poisson(const uaamg_grid* liquid_phi, const uaamg_grid* face_weights, float dt, float dx,
                    const uaamg_options* opts, PoissonHierarchyHost& hier, SolverOptions& so, bool device_resident) {
    if (!(dx > 0.f) || !(dt > 0.f)) return fail(UAAMG_EINVAL, "dt and dx must be positive");
    uaamg_options o;
    uaamg_options_default(&o);
    if (opts) o = *opts;
    if (o.smoother < 0 || o.smoother > 3) return fail(UAAMG_EINVAL, "unknown smoother");
    if (o.max_iterations < 0) return fail(UAAMG_EINVAL, "max_iterations < 0");
    if (o.precision != UAAMG_PRECISION_FP32 && o.precision != UAAMG_PRECISION_BF16_CYCLE)
        return fail(UAAMG_EINVAL, "unsupported precision");
    const bool host_only = o.device < 0;   // hierarchy inspection without a GPU; solves fail
    if (!host_only && uaamg_device_count() <= o.device) return fail(UAAMG_ENODEV, "no HIP device");
    StageClock clock;
    HostGrid phi, w;
    if (!to_host_grid(liquid_phi, 1, phi, "liquid_phi")) return UAAMG_EINVAL;
    if (!to_host_grid(face_weights, 3, w, "face_weights")) return UAAMG_EINVAL;
    clock.lap("input grids to tree order");
    // the per-voxel level-0 stage runs on the handle's device (bit-identical to the host stage)
    std::string err;
    const bool host_stage = host_only || (o.flags & 256);   // flags bit 8: the whole build on the host threads
    const int64_t max_coarsest = o.max_coarsest_dof > 0 ? o.max_coarsest_dof : 4000;
    if (device_resident && !host_stage) {
        if (!build_hierarchy_device(phi, w, dt, dx, max_coarsest, hier, o.device, err))
            return fail(UAAMG_EHIP, err.empty() ? "device hierarchy build failed" : err);
    } else if (!build_hierarchy(phi, w, dt, dx, max_coarsest, hier, host_stage ? -1 : o.device, &err)) {
        return fail(err.empty() ? UAAMG_EINVAL : UAAMG_EHIP, err.empty() ? "bad grids" : err);
    }
    clock.lap("hierarchy");
    if (hier.levels[0].ndof == 0) return fail(UAAMG_EINVAL, "no degrees of freedom");
    so = SolverOptions{};
    so.relative_tolerance = o.relative_tolerance;
    so.max_iterations = o.max_iterations;
    so.smoother = o.smoother;
    so.precision = o.precision;
    so.device = o.device;
    so.max_coarsest_dof = o.max_coarsest_dof;
    so.flags = o.flags;
    so.stream = static_cast<hipStream_t>(o.stream);
    return 0;
}
}  // namespace capi
}  // namespace uaamg

extern "C" {

int uaamg_poisson_create(const uaamg_grid* liquid_phi, const uaamg_grid* face_weights, float dt, float dx,
                         const uaamg_options* opts, uaamg_poisson** out) {
    UAAMG_GUARD_BEGIN
    if (!out) return fail(UAAMG_EINVAL, "null output handle");
    *out = nullptr;
    PoissonHierarchyHost hier;
    SolverOptions so;
    if (int rc = capi::prepare_poisson(liquid_phi, face_weights, dt, dx, opts, hier, so, true)) return rc;
    const bool host_only = so.device < 0;
    std::unique_ptr<uaamg_poisson> h(new uaamg_poisson());
    if (host_only) {
        h->dev.init_host_only(std::move(hier), so);
        h->ndof = h->dev.host().levels[0].ndof;
        *out = h.release();
        return 0;
    }
    int rc = h->dev.init(std::move(hier), so);
    if (rc < 0) return hip_status(rc, h->dev);
    h->ndof = h->dev.host().levels[0].ndof;
    for (int i = 0; i < 3; ++i) {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&h->dofbuf[i]), size_t(h->ndof) * sizeof(float));
        if (e != hipSuccess) return fail(UAAMG_ENOMEM, hipGetErrorString(e));
    }
    *out = h.release();
    return 0;
    UAAMG_GUARD_END
}

void uaamg_poisson_destroy(uaamg_poisson* h) { delete h; }

int uaamg_poisson_set_options(uaamg_poisson* h, const uaamg_options* o) {
    if (!h || !o) return fail(UAAMG_EINVAL, "null argument");
    if (o->smoother < 0 || o->smoother > 3) return fail(UAAMG_EINVAL, "unknown smoother");
    if (o->max_iterations < 0) return fail(UAAMG_EINVAL, "max_iterations < 0");
    h->dev.options().relative_tolerance = o->relative_tolerance;
    h->dev.options().max_iterations = o->max_iterations;
    h->dev.options().smoother = o->smoother;
    // precision and device placement are fixed when the handle is created
    return 0;
}

int64_t uaamg_poisson_ndof(const uaamg_poisson* h) { return h ? h->ndof : -1; }

int uaamg_poisson_level_triplets(const uaamg_poisson* h, int level, int64_t* count, int32_t* rows, int32_t* cols,
                                 float* vals) {
    UAAMG_GUARD_BEGIN
    if (!h || !count || level < 0 || level >= h->dev.nlevels()) return fail(UAAMG_EINVAL, "bad argument");
    // a query of the explicit matrix reads the host bricks (copied from the device on first use)
    if (const_cast<PoissonDevice&>(h->dev).ensure_host_coefficients()) return fail(UAAMG_EHIP, h->dev.error());
    std::vector<int32_t> r, c;
    std::vector<float> v;
    level_triplets(h->dev.host().levels[size_t(level)], r, c, v);
    *count = int64_t(r.size());
    if (rows) std::memcpy(rows, r.data(), r.size() * sizeof(int32_t));
    if (cols) std::memcpy(cols, c.data(), c.size() * sizeof(int32_t));
    if (vals) std::memcpy(vals, v.data(), v.size() * sizeof(float));
    return 0;
    UAAMG_GUARD_END
}

int uaamg_poisson_kernel_map(const uaamg_poisson* h, int level, int32_t* block, float* ms_block, float* ms_sweep) {
    if (!h || level < 0 || level >= h->dev.nlevels()) return fail(UAAMG_EINVAL, "bad level");
    const auto& m = h->dev.block_levels();
    if (block) *block = level < int(m.size()) && m[size_t(level)] ? (h->dev.post_cg_level(level) ? 2 : 1) : 0;
    if (ms_block) *ms_block = h->dev.tune_ms(level, 0);
    if (ms_sweep) *ms_sweep = h->dev.tune_ms(level, 1);
    return 0;
}

int uaamg_poisson_last_residual(const uaamg_poisson* h, float* true_residual_inf, int32_t* replacements) {
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    if (true_residual_inf) *true_residual_inf = h->dev.last_true_residual();
    if (replacements) *replacements = h->dev.last_replacements();
    return 0;
}

int uaamg_poisson_nlevels(const uaamg_poisson* h) { return h ? h->dev.nlevels() : -1; }

int uaamg_poisson_level_info(const uaamg_poisson* h, int level, int64_t* ndof, int64_t* nleaf_dof,
                             int64_t* nleaf_tree) {
    if (!h || level < 0 || level >= h->dev.nlevels()) return fail(UAAMG_EINVAL, "bad level");
    const auto& l = h->dev.host().levels[level];
    if (ndof) *ndof = l.ndof;
    if (nleaf_dof) *nleaf_dof = int64_t(l.nleaf());
    if (nleaf_tree) *nleaf_tree = int64_t(l.full_origin.size());
    return 0;
}

int uaamg_poisson_level_topology(const uaamg_poisson* h, int level, int32_t* origins, uint64_t* masks,
                                 int64_t* dofbase) {
    if (!h || level < 0 || level >= h->dev.nlevels()) return fail(UAAMG_EINVAL, "bad level");
    const auto& l = h->dev.host().levels[level];
    for (size_t i = 0; i < l.nleaf(); ++i) {
        if (origins) {
            origins[3 * i] = l.origin[i].x;
            origins[3 * i + 1] = l.origin[i].y;
            origins[3 * i + 2] = l.origin[i].z;
        }
        if (dofbase) dofbase[i] = l.dofbase[i];
    }
    if (masks && !l.mask.empty()) std::memcpy(masks, l.mask.data(), l.mask.size() * sizeof(uint64_t));
    return 0;
}

int uaamg_poisson_grid_to_vector(const uaamg_poisson* h, const uaamg_grid* grid, float* vec) {
    UAAMG_GUARD_BEGIN
    if (!h || !grid || !vec) return fail(UAAMG_EINVAL, "null argument");
    HostGrid g;
    if (!to_host_grid(grid, 1, g, "grid")) return UAAMG_EINVAL;
    const auto& l = h->dev.host().levels[0];
    for (size_t i = 0; i < l.nleaf(); ++i) {
        const int32_t gi = g.lookup.find(l.origin[i]);
        const uint64_t* m = &l.mask[i * 8];
        int64_t d = l.dofbase[i];
        for (int off = 0; off < kLeafSize; ++off)
            if (mask_test(m, off)) vec[d++] = gi < 0 ? 0.f : g.leaf_values(size_t(gi))[off];
    }
    return 0;
    UAAMG_GUARD_END
}

int uaamg_poisson_vector_to_grid(const uaamg_poisson* h, const float* vec, float* bricks) {
    UAAMG_GUARD_BEGIN
    if (!h || !vec || !bricks) return fail(UAAMG_EINVAL, "null argument");
    const auto& l = h->dev.host().levels[0];
    for (size_t i = 0; i < l.nleaf(); ++i) {
        const uint64_t* m = &l.mask[i * 8];
        float* out = bricks + i * kLeafSize;
        int64_t d = l.dofbase[i];
        for (int off = 0; off < kLeafSize; ++off) out[off] = mask_test(m, off) ? vec[d++] : 0.f;
    }
    return 0;
    UAAMG_GUARD_END
}

int uaamg_poisson_level_coefficients(const uaamg_poisson* h, int level, int kind, float* bricks, uint8_t* kept) {
    if (!h || level < 0 || level >= h->dev.nlevels()) return fail(UAAMG_EINVAL, "bad level");
    if (kind < 0 || kind >= kNumCoef) return fail(UAAMG_EINVAL, "bad coefficient kind");
    // the SPAI0 coefficients are built on first use; a query is a use (the handle's
    // logical state does not change)
    if (kind == kCoefSpai0 && const_cast<PoissonDevice&>(h->dev).ensure_spai0())
        return fail(UAAMG_EHIP, h->dev.error());
    if (bricks && const_cast<PoissonDevice&>(h->dev).ensure_host_coefficients())
        return fail(UAAMG_EHIP, h->dev.error());
    const auto& l = h->dev.host().levels[level];
    if (bricks && !l.coef[kind].empty())
        std::memcpy(bricks, l.coef[kind].data(), l.coef[kind].size() * sizeof(float));
    if (kept && !l.keep[kind].empty()) std::memcpy(kept, l.keep[kind].data(), l.keep[kind].size());
    return 0;
}

int uaamg_poisson_build_rhs(const uaamg_poisson* h, const uaamg_grid* face_weights, const uaamg_grid* vx,
                            const uaamg_grid* vy, const uaamg_grid* vz, const uaamg_grid* solid_velocity,
                            float* rhs) {
    UAAMG_GUARD_BEGIN
    if (!h || !rhs) return fail(UAAMG_EINVAL, "null argument");
    HostGrid w, gx, gy, gz, sv;
    if (!to_host_grid(face_weights, 3, w, "face_weights")) return UAAMG_EINVAL;
    if (!to_host_grid(vx, 1, gx, "vx") || !to_host_grid(vy, 1, gy, "vy") || !to_host_grid(vz, 1, gz, "vz"))
        return UAAMG_EINVAL;
    if (solid_velocity) {
        if (!to_host_grid(solid_velocity, 3, sv, "solid_velocity")) return UAAMG_EINVAL;
    } else {
        sv.ncomp = 3;
        sv.lookup.build(sv.origin);
    }
    build_poisson_rhs(h->dev.host().levels[0], w, gx, gy, gz, sv, rhs);
    return 0;
    UAAMG_GUARD_END
}

namespace {

// the five input grids of a right-hand-side build as RhsGridRef (kRhs* order);
// returns 0 or UAAMG_EINVAL with the message set
int rhs_refs(const uaamg_grid* const in[kRhsGrids], bool values_on_device, RhsGridRef refs[kRhsGrids]) {
    const int ncomp[kRhsGrids] = {3, 1, 1, 1, 3};
    const char* names[kRhsGrids] = {"face_weights", "vx", "vy", "vz", "solid_velocity"};
    for (int g = 0; g < kRhsGrids; ++g) {
        RhsGridRef& r = refs[g];
        r.ncomp = ncomp[g];
        if (!in[g]) {
            if (g != kRhsSolid) return fail(UAAMG_EINVAL, std::string(names[g]) + ": null grid");
            continue;   // zero solid velocity: an empty grid with background 0
        }
        if (in[g]->ncomp != ncomp[g]) return fail(UAAMG_EINVAL, std::string(names[g]) + ": wrong component count");
        if (in[g]->nleaf < 0 || (in[g]->nleaf > 0 && (!in[g]->origins || !in[g]->values)))
            return fail(UAAMG_EINVAL, std::string(names[g]) + ": null arrays");
        r.nleaf = in[g]->nleaf;
        r.origins = in[g]->origins;
        r.values = in[g]->values;
        r.values_on_device = values_on_device;
        for (int c = 0; c < 3; ++c) r.bg[c] = in[g]->background[c];
    }
    return 0;
}

}  // namespace

int uaamg_poisson_build_rhs_device(uaamg_poisson* h, const uaamg_grid* face_weights, const uaamg_grid* vx,
                                   const uaamg_grid* vy, const uaamg_grid* vz, const uaamg_grid* solid_velocity,
                                   int32_t values_on_device, float* rhs_dev) {
    UAAMG_GUARD_BEGIN
    if (!h || !rhs_dev) return fail(UAAMG_EINVAL, "null argument");
    const uaamg_grid* in[kRhsGrids] = {face_weights, vx, vy, vz, solid_velocity};
    RhsGridRef refs[kRhsGrids];
    if (int rc = rhs_refs(in, values_on_device != 0, refs)) return rc;
    std::string err;
    const int rc = build_poisson_rhs_device(h->dev, refs, rhs_dev, err);
    if (rc < 0) return fail(rc == -4 ? UAAMG_ENODEV : (rc == -1 ? UAAMG_EINVAL : rc), err);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_poisson_build_rhs_device_noalloc(uaamg_poisson* h, const uaamg_grid* face_weights, const uaamg_grid* vx,
                                           const uaamg_grid* vy, const uaamg_grid* vz,
                                           const uaamg_grid* solid_velocity, float* rhs_dev) {
    UAAMG_GUARD_BEGIN
    if (!h || !rhs_dev) return fail(UAAMG_EINVAL, "null argument");
    const uaamg_grid* in[kRhsGrids] = {face_weights, vx, vy, vz, solid_velocity};
    RhsGridRef refs[kRhsGrids];
    if (int rc = rhs_refs(in, true, refs)) return rc;
    std::string err;
    const int rc = build_poisson_rhs_device_noalloc(h->dev, refs, rhs_dev, err);
    if (rc < 0) return fail(rc == -4 ? UAAMG_ENODEV : (rc == -1 ? UAAMG_EINVAL : rc), err);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_poisson_apply_pressure_gradient_device(uaamg_poisson* h, const float* pressure_dev,
                                                 const uaamg_grid* liquid_phi, const uaamg_grid* face_weights,
                                                 const uaamg_grid* solid_velocity, uaamg_face_channel* channels,
                                                 float dt, float friction, int32_t values_on_device) {
    UAAMG_GUARD_BEGIN
    if (!h || !pressure_dev || !channels) return fail(UAAMG_EINVAL, "null argument");
    const uaamg_grid* in[3] = {liquid_phi, face_weights, solid_velocity};
    const int ncomp[3] = {1, 3, 3};
    const char* names[3] = {"liquid_phi", "face_weights", "solid_velocity"};
    FaceGridRef refs[3];
    for (int g = 0; g < 3; ++g) {
        FaceGridRef& r = refs[g];
        r.ncomp = ncomp[g];
        if (!in[g]) {
            if (g != 2) return fail(UAAMG_EINVAL, std::string(names[g]) + ": null grid");
            continue;   // zero solid velocity
        }
        if (in[g]->ncomp != ncomp[g]) return fail(UAAMG_EINVAL, std::string(names[g]) + ": wrong component count");
        if (in[g]->nleaf < 0 || (in[g]->nleaf > 0 && (!in[g]->origins || !in[g]->values)))
            return fail(UAAMG_EINVAL, std::string(names[g]) + ": null arrays");
        r.nleaf = in[g]->nleaf;
        r.origins = in[g]->origins;
        r.values = in[g]->values;
        for (int k = 0; k < 3; ++k) r.bg[k] = in[g]->background[k];
    }
    FaceChannelRef ch[3];
    for (int c = 0; c < 3; ++c) {
        const uaamg_face_channel& f = channels[c];
        if (f.nleaf < 0 || (f.nleaf > 0 && (!f.origins || !f.active || !f.values)))
            return fail(UAAMG_EINVAL, "velocity channel: null arrays");
        ch[c].nleaf = f.nleaf;
        ch[c].origins = f.origins;
        ch[c].active = f.active;
        ch[c].values = f.values;
        ch[c].update = f.update;
        ch[c].reflected = f.reflected;
    }
    std::string err;
    const int rc = apply_pressure_gradient_device(h->dev, pressure_dev, refs, ch, dt, friction, values_on_device != 0,
                                                  err);
    if (rc < 0) return fail(rc == -4 ? UAAMG_ENODEV : (rc == -1 ? UAAMG_EINVAL : rc), err);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_face_weights(const uaamg_grid* liquid_sdf, const uaamg_grid* solid_sdf, int32_t device, int64_t* nleaf,
                       int32_t* origins, uint64_t* active, float* values) {
    return uaamg_face_weights_ex(liquid_sdf, solid_sdf, device, UAAMG_FACE_WEIGHTS_STATIC, nleaf, origins, active,
                                 values);
}

int uaamg_face_weights_ex(const uaamg_grid* liquid_sdf, const uaamg_grid* solid_sdf, int32_t device, int32_t variant,
                          int64_t* nleaf, int32_t* origins, uint64_t* active, float* values) {
    UAAMG_GUARD_BEGIN
    if (!nleaf) return fail(UAAMG_EINVAL, "null argument");
    if (variant != UAAMG_FACE_WEIGHTS_STATIC && variant != UAAMG_FACE_WEIGHTS_MEMBER)
        return fail(UAAMG_EINVAL, "unknown face-weight variant");
    HostGrid liquid;
    if (!to_host_grid(liquid_sdf, 1, liquid, "liquid_sdf")) return UAAMG_EINVAL;
    if (!solid_sdf || solid_sdf->ncomp != 1 || solid_sdf->nleaf < 0 ||
        (solid_sdf->nleaf > 0 && (!solid_sdf->origins || !solid_sdf->values)))
        return fail(UAAMG_EINVAL, "solid_sdf: null grid, arrays or wrong component count");
    const bool fill = origins && active && values;
    std::vector<Coord> org;
    std::vector<uint64_t> act;
    std::string err;
    const float cut = variant == UAAMG_FACE_WEIGHTS_MEMBER ? 1e-2f : 0.f;
    const int rc = face_weights_device(liquid, solid_sdf->nleaf, solid_sdf->origins, solid_sdf->values,
                                       solid_sdf->background[0], cut, device, org, act, fill ? values : nullptr, err);
    if (rc < 0) return fail(rc == -4 ? UAAMG_ENODEV : (rc == -1 ? UAAMG_EINVAL : rc), err);
    *nleaf = int64_t(org.size());
    if (fill) {
        for (size_t i = 0; i < org.size(); ++i) {
            origins[3 * i] = org[i].x;
            origins[3 * i + 1] = org[i].y;
            origins[3 * i + 2] = org[i].z;
        }
        std::memcpy(active, act.data(), act.size() * sizeof(uint64_t));
    }
    return 0;
    UAAMG_GUARD_END
}

int uaamg_poisson_solve_pcg_device(uaamg_poisson* h, float* x_dev, const float* b_dev, int32_t* iterations) {
    if (!h || !x_dev || !b_dev) return fail(UAAMG_EINVAL, "null argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    int it = 0;
    int rc = h->dev.solve_pcg_dev(x_dev, b_dev, &it);
    if (iterations) *iterations = it;
    return hip_status(rc, h->dev);
}

int uaamg_poisson_solve_batch_device(uaamg_poisson* const* hs, int32_t n, float* const* x_dev,
                                     const float* const* b_dev, int32_t* status, int32_t* iterations,
                                     int32_t concurrency) {
    UAAMG_GUARD_BEGIN
    if (n < 0 || (n > 0 && (!hs || !x_dev || !b_dev || !status))) return fail(UAAMG_EINVAL, "null argument");
    for (int32_t i = 0; i < n; ++i) {
        if (!hs[i] || !x_dev[i] || !b_dev[i]) return fail(UAAMG_EINVAL, "null handle or vector in the batch");
        if (!hs[i]->dev.on_device()) return fail(UAAMG_ENODEV, "a batch handle was created without a device");
        for (int32_t j = 0; j < i; ++j)
            if (hs[j] == hs[i]) return fail(UAAMG_EINVAL, "a handle appears twice in the batch");
    }
    const int nt = std::max(1, std::min<int>(std::min<int>(concurrency, 16), n));
    std::atomic<int32_t> next{0};
    std::vector<std::string> errs(size_t(std::max(n, 1)));
    auto worker = [&]() {
        for (;;) {
            const int32_t i = next.fetch_add(1);
            if (i >= n) return;
            PoissonDevice& d = hs[i]->dev;
            int it = 0;
            int rc = hipSetDevice(d.options().device) == hipSuccess ? d.solve_pcg_dev(x_dev[i], b_dev[i], &it) : -3;
            status[i] = rc;
            if (iterations) iterations[i] = it;
            if (rc < 0) errs[size_t(i)] = d.error().empty() ? "device error" : d.error();
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    for (int32_t i = 0; i < n; ++i)
        if (status[i] < 0) return fail(UAAMG_EHIP, "batch problem " + std::to_string(i) + ": " + errs[size_t(i)]);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_poisson_solve_pcg(uaamg_poisson* h, float* x, const float* b, int32_t* iterations) {
    if (!h || !x || !b) return fail(UAAMG_EINVAL, "null argument");
    if (int rc = copy_in(h, 0, x, h->ndof)) return rc;
    if (int rc = copy_in(h, 1, b, h->ndof)) return rc;
    int it = 0;
    int rc = h->dev.solve_pcg_dev(h->dofbuf[0], h->dofbuf[1], &it);
    if (iterations) *iterations = it;
    if (rc < 0) return hip_status(rc, h->dev);
    if (int rc2 = copy_out(h, 0, x, h->ndof)) return rc2;
    return rc;
}

int uaamg_poisson_solve_mg(uaamg_poisson* h, float* x, const float* b, int32_t* iterations) {
    if (!h || !x || !b) return fail(UAAMG_EINVAL, "null argument");
    if (int rc = copy_in(h, 0, x, h->ndof)) return rc;
    if (int rc = copy_in(h, 1, b, h->ndof)) return rc;
    int it = 0;
    int rc = h->dev.solve_mg_dev(h->dofbuf[0], h->dofbuf[1], &it);
    if (iterations) *iterations = it;
    if (rc < 0) return hip_status(rc, h->dev);
    if (int rc2 = copy_out(h, 0, x, h->ndof)) return rc2;
    return rc;
}

int uaamg_poisson_apply(uaamg_poisson* h, int level, int mode, float weight, const float* x, const float* b,
                        float* y) {
    if (!h || !x || !y || level < 0 || level >= h->dev.nlevels()) return fail(UAAMG_EINVAL, "bad argument");
    if (mode < 0 || mode > 5) return fail(UAAMG_EINVAL, "bad mode");
    if (mode != UAAMG_MODE_LAPLACIAN && !b) return fail(UAAMG_EINVAL, "mode needs b");
    const int64_t n = level_ndof(h, level);
    if (int rc = copy_in(h, 0, x, n)) return rc;
    if (int rc = copy_in(h, 1, b, n)) return rc;
    int rc = h->dev.apply_dev(level, mode, weight, h->dofbuf[0], b ? h->dofbuf[1] : nullptr, h->dofbuf[2]);
    if (rc < 0) return hip_status(rc, h->dev);
    return copy_out(h, 2, y, n);
}

int uaamg_poisson_first_sweep(uaamg_poisson* h, int level, int kind, float weight, const float* b, float* y) {
    if (!h || !b || !y || level < 0 || level >= h->dev.nlevels() || kind < 0 || kind > 2)
        return fail(UAAMG_EINVAL, "bad argument");
    const int64_t n = level_ndof(h, level);
    if (int rc = copy_in(h, 1, b, n)) return rc;
    int rc = h->dev.first_sweep_dev(level, kind, weight, h->dofbuf[1], h->dofbuf[2]);
    if (rc < 0) return hip_status(rc, h->dev);
    return copy_out(h, 2, y, n);
}

int uaamg_poisson_restrict(uaamg_poisson* h, int fine_level, const float* fine, float* coarse) {
    if (!h || !fine || !coarse || fine_level < 0 || fine_level + 1 >= h->dev.nlevels())
        return fail(UAAMG_EINVAL, "bad argument");
    if (int rc = copy_in(h, 0, fine, level_ndof(h, fine_level))) return rc;
    int rc = h->dev.restrict_dev(fine_level, h->dofbuf[0], h->dofbuf[2]);
    if (rc < 0) return hip_status(rc, h->dev);
    return copy_out(h, 2, coarse, level_ndof(h, fine_level + 1));
}

int uaamg_poisson_prolong(uaamg_poisson* h, int fine_level, const float* coarse, float* fine, float alpha) {
    if (!h || !fine || !coarse || fine_level < 0 || fine_level + 1 >= h->dev.nlevels())
        return fail(UAAMG_EINVAL, "bad argument");
    if (int rc = copy_in(h, 1, coarse, level_ndof(h, fine_level + 1))) return rc;
    if (int rc = copy_in(h, 0, fine, level_ndof(h, fine_level))) return rc;
    int rc = h->dev.prolong_dev(fine_level, h->dofbuf[1], h->dofbuf[0], alpha);
    if (rc < 0) return hip_status(rc, h->dev);
    return copy_out(h, 0, fine, level_ndof(h, fine_level));
}

int uaamg_poisson_precondition(uaamg_poisson* h, const float* r, float* z) {
    if (!h || !r || !z) return fail(UAAMG_EINVAL, "null argument");
    if (int rc = copy_in(h, 1, r, h->ndof)) return rc;
    int rc = h->dev.precondition_dev(h->dofbuf[1], h->dofbuf[2]);
    if (rc < 0) return hip_status(rc, h->dev);
    return copy_out(h, 2, z, h->ndof);
}

int uaamg_poisson_coarsest_solve(uaamg_poisson* h, const float* b, float* x) {
    if (!h || !b || !x) return fail(UAAMG_EINVAL, "null argument");
    const int l = h->dev.nlevels() - 1;
    const int64_t n = level_ndof(h, l);
    if (int rc = copy_in(h, 1, b, n)) return rc;
    int rc = h->dev.coarsest_dev(h->dofbuf[1], h->dofbuf[2]);
    if (rc < 0) return hip_status(rc, h->dev);
    return copy_out(h, 2, x, n);
}

int uaamg_poisson_algorithmic_bytes(const uaamg_poisson* h, int level, int mode, double* bytes) {
    if (!h || !bytes || level < 0 || level >= h->dev.nlevels() || mode < 0 || mode > UAAMG_MODE_COARSEST)
        return fail(UAAMG_EINVAL, "bad argument");
    *bytes = h->dev.algorithmic_bytes(level, mode);
    return 0;
}

int uaamg_poisson_time_apply(uaamg_poisson* h, int level, int mode, int reps, float* ms_per_launch) {
    if (!h || !ms_per_launch) return fail(UAAMG_EINVAL, "null argument");
    if (!h->dev.on_device()) return fail(UAAMG_ENODEV, "handle was created without a device (device < 0)");
    int rc = h->dev.time_apply(level, mode, reps, ms_per_launch);
    if (rc == -1) return fail(UAAMG_EINVAL, "bad argument");
    return hip_status(rc, h->dev);
}

// ------------------------------------------------------------------ scenes
int uaamg_scene_create(int config, int32_t n, uint64_t seed, uaamg_scene** out) {
    UAAMG_GUARD_BEGIN
    if (!out) return fail(UAAMG_EINVAL, "null output");
    SceneSpec spec;
    if (!scene_preset(config, n, seed, spec)) return fail(UAAMG_EINVAL, "unknown scene config");
    std::unique_ptr<uaamg_scene> s(new uaamg_scene());
    build_scene(spec, s->s);
    *out = s.release();
    return 0;
    UAAMG_GUARD_END
}

int uaamg_scene_create_custom(int32_t n, float pool, int32_t nspheres, const float* spheres, int32_t has_solid,
                              const float* solid_sphere, uint64_t velocity_seed, uaamg_scene** out) {
    UAAMG_GUARD_BEGIN
    if (!out || n <= 0 || nspheres < 0 || (nspheres > 0 && !spheres)) return fail(UAAMG_EINVAL, "bad argument");
    SceneSpec spec;
    spec.n = n;
    spec.pool = pool;
    spec.spheres.assign(spheres, spheres + 4 * nspheres);
    spec.has_solid = has_solid != 0;
    if (solid_sphere)
        for (int i = 0; i < 4; ++i) spec.solid[i] = solid_sphere[i];
    spec.velocity_seed = velocity_seed;
    std::unique_ptr<uaamg_scene> s(new uaamg_scene());
    build_scene(spec, s->s);
    *out = s.release();
    return 0;
    UAAMG_GUARD_END
}

int uaamg_scene_grids(const uaamg_scene* sc, uaamg_grid* phi, uaamg_grid* weights, uaamg_grid* vx, uaamg_grid* vy,
                      uaamg_grid* vz, float* dx, int32_t* n) {
    if (!sc) return fail(UAAMG_EINVAL, "null scene");
    const Scene& s = sc->s;
    auto fill = [&](uaamg_grid* g, const float* values, int ncomp, float bg) {
        if (!g) return;
        g->nleaf = s.nleaf();
        g->origins = s.origins.data();
        g->active = s.active.data();
        g->values = values;
        g->ncomp = ncomp;
        g->background[0] = g->background[1] = g->background[2] = bg;
    };
    fill(phi, s.phi.data(), 1, s.phi_background);
    fill(weights, s.weights.data(), 3, 0.f);
    fill(vx, s.vel[0].data(), 1, 0.f);
    fill(vy, s.vel[1].data(), 1, 0.f);
    fill(vz, s.vel[2].data(), 1, 0.f);
    if (dx) *dx = s.dx;
    if (n) *n = s.spec.n;
    return 0;
}

void uaamg_scene_destroy(uaamg_scene* s) { delete s; }

float uaamg_fraction_inside2(float a, float b) { return fraction_inside(a, b); }
float uaamg_fraction_inside4(float bl, float br, float tl, float tr) { return fraction_inside(bl, br, tl, tr); }

}  // extern "C"

// ------------------------------------------------------------ merged batch

struct uaamg_poisson_batch {
    PoissonBatch b;
};

extern "C" {

int uaamg_poisson_batch_create(int32_t n, const uaamg_grid* const* liquid_phi, const uaamg_grid* const* face_weights,
                               const float* dt, const float* dx, const uaamg_options* opts,
                               uaamg_poisson_batch** out) {
    UAAMG_GUARD_BEGIN
    if (!out) return fail(UAAMG_EINVAL, "null output handle");
    *out = nullptr;
    if (n <= 0 || !liquid_phi || !face_weights || !dt || !dx) return fail(UAAMG_EINVAL, "empty or null batch");
    std::vector<PoissonHierarchyHost> probs(static_cast<size_t>(n));
    SolverOptions so;
    for (int32_t i = 0; i < n; ++i) {
        SolverOptions oi;
        if (int rc = capi::prepare_poisson(liquid_phi[i], face_weights[i], dt[i], dx[i], opts, probs[size_t(i)], oi))
            return rc;
        if (i == 0) so = oi;
    }
    if (so.device < 0) return fail(UAAMG_ENODEV, "the merged batch needs a device");
    std::unique_ptr<uaamg_poisson_batch> h(new uaamg_poisson_batch());
    int rc = h->b.init(std::move(probs), so);
    if (rc < 0) {
        const int code = rc == -1 ? UAAMG_EINVAL : rc == -2 ? UAAMG_ENOMEM : rc == -4 ? UAAMG_ENODEV : UAAMG_EHIP;
        return fail(code, h->b.error().empty() ? "batch initialisation failed" : h->b.error());
    }
    *out = h.release();
    return 0;
    UAAMG_GUARD_END
}

void uaamg_poisson_batch_destroy(uaamg_poisson_batch* h) { delete h; }

int uaamg_poisson_batch_set_options(uaamg_poisson_batch* h, const uaamg_options* o) {
    if (!h || !o) return fail(UAAMG_EINVAL, "null argument");
    if (o->max_iterations < 0) return fail(UAAMG_EINVAL, "max_iterations < 0");
    h->b.options().relative_tolerance = o->relative_tolerance;
    h->b.options().max_iterations = o->max_iterations;
    return 0;
}

int uaamg_poisson_batch_info(const uaamg_poisson_batch* h, int32_t* nproblems, int32_t* nlevels, int64_t* leaves0,
                             int64_t* device_bytes) {
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    if (nproblems) *nproblems = h->b.nproblems();
    if (nlevels) *nlevels = h->b.nlevels();
    if (leaves0) *leaves0 = h->b.level_leaves(0);
    if (device_bytes) *device_bytes = int64_t(h->b.device_bytes());
    return 0;
}

int64_t uaamg_poisson_batch_ndof(const uaamg_poisson_batch* h, int32_t i) { return h ? h->b.ndof(i) : -1; }

int uaamg_poisson_batch_last_residual(const uaamg_poisson_batch* h, float* true_residual_inf, int32_t* replacements) {
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    for (int i = 0; i < h->b.nproblems(); ++i) {
        if (true_residual_inf) true_residual_inf[i] = h->b.last_true_residual(i);
        if (replacements) replacements[i] = h->b.last_replacements(i);
    }
    return 0;
}

int uaamg_poisson_batch_last_fallback(const uaamg_poisson_batch* h, int32_t* mg_status, int32_t* mg_iterations) {
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    for (int i = 0; i < h->b.nproblems(); ++i) {
        if (mg_status) mg_status[i] = h->b.last_mg_status(i);
        if (mg_iterations) mg_iterations[i] = h->b.last_mg_iterations(i);
    }
    return 0;
}

int uaamg_poisson_batch_solve_device(uaamg_poisson_batch* h, float* const* x_dev, const float* const* b_dev,
                                     int32_t* status, int32_t* iterations) {
    UAAMG_GUARD_BEGIN
    if (!h || !x_dev || !b_dev) return fail(UAAMG_EINVAL, "null argument");
    const int n = h->b.nproblems();
    std::vector<int> st(static_cast<size_t>(n)), it(static_cast<size_t>(n));
    int rc = h->b.solve(x_dev, b_dev, st.data(), it.data());
    if (rc < 0) {
        const int code = rc == -1 ? UAAMG_EINVAL : rc == -4 ? UAAMG_ENODEV : UAAMG_EHIP;
        return fail(code, h->b.error().empty() ? "batch solve failed" : h->b.error());
    }
    for (int i = 0; i < n; ++i) {
        if (status) status[i] = st[size_t(i)];
        if (iterations) iterations[i] = it[size_t(i)];
    }
    return 0;
    UAAMG_GUARD_END
}

}  // extern "C"
