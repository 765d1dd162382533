// C ABI of the solve server (include/uaamg_capi.h, server section).
#include <hip/hip_runtime.h>

#include <memory>
#include <string>

#include "capi_common.h"
#include "server.h"

using namespace uaamg;
using uaamg::capi::fail;

struct uaamg_server {
    std::unique_ptr<Server> s;
};

namespace {
int server_status(int rc, const Server& s) {
    if (rc == 0) return 0;
    const int code = rc == -1 ? UAAMG_EINVAL : rc == -2 ? UAAMG_ENOMEM : rc == -4 ? UAAMG_ENODEV : UAAMG_EHIP;
    return fail(code, s.error().empty() ? "server error" : s.error());
}
}  // namespace

extern "C" {

int uaamg_server_options_default(uaamg_server_options* o) {
    if (!o) return fail(UAAMG_EINVAL, "null options");
    o->device = 0;
    o->slots = 4;
    o->memory_budget_bytes = 0;
    return 0;
}

int uaamg_server_create(const uaamg_server_options* opts, uaamg_server** out) {
    UAAMG_GUARD_BEGIN
    if (!out) return fail(UAAMG_EINVAL, "null output handle");
    *out = nullptr;
    uaamg_server_options o;
    uaamg_server_options_default(&o);
    if (opts) o = *opts;
    int ndev = 0;
    if (o.device < 0 || hipGetDeviceCount(&ndev) != hipSuccess || ndev <= o.device)
        return fail(UAAMG_ENODEV, "no HIP device with that ordinal");
    ServerOptions so;
    so.device = o.device;
    so.slots = o.slots;
    so.memory_budget = o.memory_budget_bytes;
    std::unique_ptr<uaamg_server> h(new uaamg_server());
    h->s.reset(new Server(so));
    if (int rc = h->s->start()) return server_status(rc, *h->s);
    *out = h.release();
    return 0;
    UAAMG_GUARD_END
}

void uaamg_server_destroy(uaamg_server* h) { delete h; }

int uaamg_server_add_poisson(uaamg_server* h, const uaamg_grid* liquid_phi, const uaamg_grid* face_weights, float dt,
                             float dx, const uaamg_options* opts, int64_t* problem) {
    UAAMG_GUARD_BEGIN
    if (!h || !problem) return fail(UAAMG_EINVAL, "null argument");
    PoissonHierarchyHost hier;
    SolverOptions so;
    if (int rc = capi::prepare_poisson(liquid_phi, face_weights, dt, dx, opts, hier, so)) return rc;
    *problem = h->s->add_poisson(std::move(hier), so);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_server_add_viscosity(uaamg_server* h, const uaamg_grid* viscosity, const uaamg_grid* liquid_sdf,
                               const uaamg_grid* solid_sdf, float dt, float rho, float dx,
                               const uaamg_viscosity_options* opts, int64_t* problem) {
    UAAMG_GUARD_BEGIN
    if (!h || !problem) return fail(UAAMG_EINVAL, "null argument");
    ViscHierarchyHost hh;
    ViscOptions vo;
    if (int rc = capi::prepare_viscosity(viscosity, liquid_sdf, solid_sdf, dt, rho, dx, opts, hh, vo)) return rc;
    *problem = h->s->add_viscosity(std::move(hh), vo);
    return 0;
    UAAMG_GUARD_END
}

int uaamg_server_problem_info(uaamg_server* h, int64_t problem, int64_t* ndof, int64_t* device_bytes,
                              int32_t* resident) {
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    int res = 0;
    if (!h->s->problem_info(problem, ndof, device_bytes, &res)) return fail(UAAMG_EINVAL, "unknown problem id");
    if (resident) *resident = res;
    return 0;
}

int uaamg_server_submit(uaamg_server* h, int64_t problem, const float* b_dev, float* x_out_dev, const float* x0_dev,
                        float tolerance, int64_t* ticket) {
    UAAMG_GUARD_BEGIN
    if (!h || !ticket) return fail(UAAMG_EINVAL, "null argument");
    if (!(tolerance > 0.f)) return fail(UAAMG_EINVAL, "tolerance must be positive");
    return server_status(h->s->submit(problem, b_dev, x_out_dev, x0_dev, tolerance, ticket), *h->s);
    UAAMG_GUARD_END
}

int uaamg_server_wait(uaamg_server* h, int64_t ticket, int32_t* status, int32_t* iterations, double* queue_ms,
                      double* total_ms) {
    UAAMG_GUARD_BEGIN
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    int st = 0, it = 0;
    std::string msg;
    if (int rc = h->s->wait(ticket, &st, &it, queue_ms, total_ms, nullptr, nullptr, &msg))
        return server_status(rc, *h->s);
    if (st < 0) (void)fail(st, msg);   // uaamg_last_error() names the failed step
    if (status) *status = st;
    if (iterations) *iterations = it;
    return 0;
    UAAMG_GUARD_END
}

int uaamg_server_wait_fallback(uaamg_server* h, int64_t ticket, int32_t* status, int32_t* iterations,
                               int32_t* mg_status, int32_t* mg_iterations, double* queue_ms, double* total_ms) {
    UAAMG_GUARD_BEGIN
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    int st = 0, it = 0, ms = -1, mi = -1;
    std::string msg;
    if (int rc = h->s->wait(ticket, &st, &it, queue_ms, total_ms, &ms, &mi, &msg)) return server_status(rc, *h->s);
    if (st < 0) (void)fail(st, msg);
    if (status) *status = st;
    if (iterations) *iterations = it;
    if (mg_status) *mg_status = ms;
    if (mg_iterations) *mg_iterations = mi;
    return 0;
    UAAMG_GUARD_END
}

int uaamg_server_stats(uaamg_server* h, uaamg_server_counters* out) {
    if (!h || !out) return fail(UAAMG_EINVAL, "null argument");
    const ServerStats s = h->s->stats();
    out->completed = s.completed;
    out->failed = s.failed;
    out->evictions = s.evictions;
    out->acquisitions = s.acquisitions;
    out->resident_bytes = s.resident_bytes;
    out->peak_inflight = s.peak_inflight;
    out->queued = s.queued;
    return 0;
}

}  // extern "C"
