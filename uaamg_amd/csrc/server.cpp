// Solve server (see server.h).
#include "server.h"

#include <algorithm>

namespace uaamg {

namespace {
constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
inline double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}
}  // namespace

Server::Server(const ServerOptions& o) : opt_(o) {}

int Server::start() {
    if (opt_.slots < 1 || opt_.slots > 16) {
        err_ = "slots must lie in [1, 16]";
        return -1;
    }
    if (hipSetDevice(opt_.device) != hipSuccess) {
        err_ = "hipSetDevice failed";
        return -4;
    }
    budget_ = size_t(opt_.memory_budget);
    if (opt_.memory_budget <= 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            err_ = "hipMemGetInfo failed";
            return -3;
        }
        budget_ = free_b / 10 * 9;
    }
    for (int i = 0; i < opt_.slots; ++i) workers_.emplace_back([this] { worker(); });
    return 0;
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
    (void)hipSetDevice(opt_.device);
    for (auto& kv : problems_) {
        Problem& p = *kv.second;
        if (p.x) (void)hipFree(p.x);
        p.x = nullptr;
    }
    problems_.clear();   // the solvers release their device state
}

int64_t Server::add_poisson(PoissonHierarchyHost&& h, const SolverOptions& so) {
    std::unique_ptr<Problem> p(new Problem());
    p->pd.reset(new PoissonDevice());
    SolverOptions o = so;
    o.device = opt_.device;
    o.stream = nullptr;   // each problem owns its stream
    p->ndof = h.levels.empty() ? 0 : h.levels[0].ndof;
    p->pd->init_host_only(std::move(h), o);
    p->bytes = p->pd->device_bytes() + align_up(size_t(p->ndof) * sizeof(float));
    std::lock_guard<std::mutex> lk(mu_);
    const int64_t id = next_pid_++;
    problems_[id] = std::move(p);
    return id;
}

int64_t Server::add_viscosity(ViscHierarchyHost&& h, const ViscOptions& vo) {
    std::unique_ptr<Problem> p(new Problem());
    p->visc = true;
    p->vd.reset(new ViscDevice());
    ViscOptions o = vo;
    o.device = opt_.device;
    o.stream = nullptr;
    p->ndof = h.levels.empty() ? 0 : h.levels[0].ndof;
    p->vd->init_host_only(std::move(h), o);
    p->bytes = p->vd->device_bytes() + align_up(size_t(p->ndof) * sizeof(float));
    std::lock_guard<std::mutex> lk(mu_);
    const int64_t id = next_pid_++;
    problems_[id] = std::move(p);
    return id;
}

bool Server::problem_info(int64_t pid, int64_t* ndof, int64_t* bytes, int* resident) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = problems_.find(pid);
    if (it == problems_.end()) return false;
    if (ndof) *ndof = it->second->ndof;
    if (bytes) *bytes = int64_t(it->second->bytes);
    if (resident) *resident = it->second->resident ? 1 : 0;
    return true;
}

int Server::submit(int64_t pid, const float* b, float* x_out, const float* x0, float tolerance, int64_t* ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = problems_.find(pid);
    if (it == problems_.end()) {
        err_ = "unknown problem id";
        return -1;
    }
    if (!b) {
        err_ = "null right-hand side";
        return -1;
    }
    if (it->second->bytes > budget_) {
        err_ = "the problem's device footprint exceeds the server's memory budget";
        return -2;
    }
    if (stop_) {
        err_ = "server is stopping";
        return -1;
    }
    Request r{next_ticket_++, pid, b, x_out, x0, tolerance, clock::now()};
    results_[r.ticket] = Result{};
    queue_.push_back(r);
    if (ticket) *ticket = r.ticket;
    cv_.notify_all();
    return 0;
}

int Server::wait(int64_t ticket, int* status, int* iterations, double* queue_ms, double* total_ms, int* mg_status,
                 int* mg_iterations, std::string* error) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = results_.find(ticket);
    if (it == results_.end()) {
        err_ = "unknown ticket";
        return -1;
    }
    cv_.wait(lk, [&] { return it->second.done; });
    const Result r = it->second;
    results_.erase(it);
    if (status) *status = r.status;
    if (iterations) *iterations = r.iterations;
    if (queue_ms) *queue_ms = r.queue_ms;
    if (total_ms) *total_ms = r.total_ms;
    if (mg_status) *mg_status = r.mg_status;
    if (mg_iterations) *mg_iterations = r.mg_iterations;
    if (error) *error = r.error;
    return 0;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ServerStats s = st_;
    s.resident_bytes = int64_t(resident_bytes_);
    s.queued = int64_t(queue_.size());
    return s;
}

// Idle resident problems to evict, least recently used first, so that `need`
// more bytes fit the budget. Called with the lock held.
bool Server::plan_room(size_t need, std::vector<Problem*>& victims) {
    victims.clear();
    if (need > budget_) return false;
    size_t avail = budget_ - std::min(budget_, resident_bytes_);
    if (need <= avail) return true;
    std::vector<Problem*> idle;
    for (auto& kv : problems_)
        if (kv.second->resident && !kv.second->busy) idle.push_back(kv.second.get());
    std::sort(idle.begin(), idle.end(), [](const Problem* a, const Problem* b) { return a->last_use < b->last_use; });
    for (Problem* v : idle) {
        if (need <= avail) break;
        victims.push_back(v);
        avail += v->bytes;
    }
    if (need <= avail) return true;
    victims.clear();
    return false;
}

void Server::evict(Problem& p) {
    // frees synchronise the device: kept apart from other workers' graph captures
    std::lock_guard<std::recursive_mutex> lk(capture_lock());
    if (p.x) {
        if (p.has_solution) {
            // on the problem's own stream (the null stream must stay clear of captures)
            hipStream_t s = p.visc ? p.vd->stream() : p.pd->stream();
            p.x_host.resize(size_t(p.ndof));
            if (hipMemcpyAsync(p.x_host.data(), p.x, size_t(p.ndof) * sizeof(float), hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                p.has_solution = false;   // restart from zero rather than from a partial copy
        }
        (void)hipFree(p.x);
        p.x = nullptr;
    }
    if (p.visc) p.vd->release_device();
    else p.pd->release_device();
}

namespace {
int hip_step(hipError_t e, const char* what, std::string& msg) {
    if (e == hipSuccess) return 0;
    msg = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? -2 : -3;
}
}  // namespace

int Server::make_resident(Problem& p, std::string& msg) {
    int rc = p.visc ? p.vd->acquire_device() : p.pd->acquire_device();
    if (rc < 0) {
        msg = "make resident: " + (p.visc ? p.vd->error() : p.pd->error());
        return rc;
    }
    const size_t nb = std::max<size_t>(size_t(p.ndof), 1) * sizeof(float);
    hipError_t me;
    {
        std::lock_guard<std::recursive_mutex> lk(capture_lock());
        me = hipMalloc(reinterpret_cast<void**>(&p.x), nb);
    }
    if (int e = hip_step(me, "make resident: solution hipMalloc", msg)) {
        p.x = nullptr;
        return e;
    }
    hipStream_t s = p.visc ? p.vd->stream() : p.pd->stream();
    hipError_t e = p.has_solution ? hipMemcpyAsync(p.x, p.x_host.data(), nb, hipMemcpyHostToDevice, s)
                                  : hipMemsetAsync(p.x, 0, nb, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return hip_step(e, "make resident: restore the solution", msg);
}

int Server::run(Problem& p, const Request& r, int* iterations, int* mg_status, int* mg_iterations,
                std::string& msg) {
    const size_t nb = size_t(p.ndof) * sizeof(float);
    hipStream_t s = p.visc ? p.vd->stream() : p.pd->stream();
    auto solver_error = [&](const char* what) {
        msg = std::string(what) + ": " + (p.visc ? p.vd->error() : p.pd->error());
    };
    // the caller's x0 and b may still be in production on its null stream:
    // the copy of x0 below is the first read of either
    if ((p.visc ? p.vd->follow_caller() : p.pd->follow_caller()) != 0) {
        solver_error("order after the caller's stream");
        return -3;
    }
    if (r.x0 && nb)
        if (int e = hip_step(hipMemcpyAsync(p.x, r.x0, nb, hipMemcpyDeviceToDevice, s), "copy x0", msg)) return e;
    int rc;
    if (p.visc) {
        p.vd->options().tolerance = r.tol;
        rc = p.vd->solve_pcg_dev(p.x, r.b, iterations);
    } else {
        p.pd->options().relative_tolerance = r.tol;
        rc = p.pd->solve_pcg_dev(p.x, r.b, iterations);
        if (rc == PoissonDevice::kFailed && (p.pd->options().flags & 128)) {
            // FLIP_vdb.cpp:3150-3155: zero the pressure and solve by pure
            // multigrid; the request keeps the PCG's FAILED status
            if (int e = hip_step(hipMemsetAsync(p.x, 0, nb, s), "zero the pressure", msg)) return e;
            const int mg = p.pd->solve_mg_dev(p.x, r.b, mg_iterations);
            if (mg < 0) {
                solver_error("pure multigrid");
                return mg;
            }
            *mg_status = mg;
        }
    }
    if (rc < 0) {
        solver_error("solve");
        return rc;
    }
    p.has_solution = true;
    if (r.x_out && nb) {
        if (int e = hip_step(hipMemcpyAsync(r.x_out, p.x, nb, hipMemcpyDeviceToDevice, s), "copy x out", msg))
            return e;
        if (int e = hip_step(hipStreamSynchronize(s), "synchronize", msg)) return e;
    }
    return rc;
}

void Server::worker() {
    (void)hipSetDevice(opt_.device);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        // the oldest request whose problem is idle and fits (after evictions)
        Problem* p = nullptr;
        std::vector<Problem*> victims;
        auto pick = queue_.end();
        for (auto q = queue_.begin(); q != queue_.end(); ++q) {
            Problem& c = *problems_[q->pid];
            if (c.busy) continue;
            if (!c.resident && !plan_room(c.bytes, victims)) continue;
            pick = q;
            p = &c;
            break;
        }
        if (!p) {
            if (stop_ && queue_.empty()) return;
            cv_.wait(lk);
            continue;
        }
        const Request r = *pick;
        queue_.erase(pick);
        p->busy = true;
        for (Problem* v : victims) {
            v->busy = true;
            v->resident = false;
            resident_bytes_ -= v->bytes;
        }
        const bool acquire = !p->resident;
        if (acquire) {
            resident_bytes_ += p->bytes;
            p->resident = true;
        }
        st_.peak_inflight = std::max(st_.peak_inflight, ++inflight_);
        const auto t_start = clock::now();
        lk.unlock();

        for (Problem* v : victims) evict(*v);
        std::string msg;
        int rc = acquire ? make_resident(*p, msg) : 0;
        const bool acquire_failed = acquire && rc != 0;
        if (acquire_failed) evict(*p);   // back to host-only
        int its = 0, mg_st = -1, mg_its = -1;
        if (rc == 0) rc = run(*p, r, &its, &mg_st, &mg_its, msg);
        const auto t_end = clock::now();

        lk.lock();
        for (Problem* v : victims) v->busy = false;
        st_.evictions += int64_t(victims.size());
        if (acquire_failed) {
            resident_bytes_ -= p->bytes;
            p->resident = false;
        } else if (acquire) {
            ++st_.acquisitions;
        }
        p->busy = false;
        p->last_use = ++use_clock_;
        --inflight_;
        Result& res = results_[r.ticket];
        res.done = true;
        res.status = rc;
        res.iterations = its;
        res.mg_status = mg_st;
        res.mg_iterations = mg_st < 0 ? -1 : mg_its;
        if (rc < 0) res.error = msg.empty() ? "server error" : msg;
        res.queue_ms = ms_between(r.submitted, t_start);
        res.total_ms = ms_between(r.submitted, t_end);
        ++st_.completed;
        if (rc != 0) ++st_.failed;
        cv_.notify_all();
    }
}

}  // namespace uaamg
