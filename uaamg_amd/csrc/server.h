// Solve server: a C++ request loop over many resident problems on one GPU.
//
// Problems (Poisson or viscosity systems of any size) are registered once;
// the server keeps their host hierarchies and makes a problem resident in
// HBM (hierarchy, PCG vectors and its warm-start solution) when one of its
// requests is admitted. Admission is continuous: each of `slots` worker
// threads takes the oldest queued request whose problem is idle and whose
// device footprint fits the memory budget, evicting idle resident problems in
// least-recently-used order when it must (an evicted problem's solution is
// kept on the host and restored with it). Every problem owns its HIP stream,
// so admitted solves of different problems overlap on the device; the solves
// of one problem run in submission order. The footprint estimate is the
// solver's arena size, a sum over the levels of per-leaf bytes times the
// level's active-leaf count.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "poisson_solver.h"
#include "viscosity_solver.h"

namespace uaamg {

struct ServerOptions {
    int device = 0;
    int slots = 4;                // worker threads = solves in flight at most
    int64_t memory_budget = 0;    // bytes of HBM for resident problems; 0: 90% of free memory at creation
};

struct ServerStats {
    int64_t completed = 0;
    int64_t failed = 0;           // completed with a FAILED or error status
    int64_t evictions = 0;
    int64_t acquisitions = 0;     // problems made resident (first time or after eviction)
    int64_t resident_bytes = 0;
    int64_t peak_inflight = 0;
    int64_t queued = 0;
};

class Server {
public:
    explicit Server(const ServerOptions& o);
    ~Server();   // finishes every queued request, then stops the workers
    int start();   // 0 or a negative uaamg status

    // registration (hierarchy already built on the host); returns the problem id
    int64_t add_poisson(PoissonHierarchyHost&& h, const SolverOptions& so);
    int64_t add_viscosity(ViscHierarchyHost&& h, const ViscOptions& vo);
    // ndof, device footprint, residency
    bool problem_info(int64_t pid, int64_t* ndof, int64_t* bytes, int* resident);

    // Queues one solve. b: right-hand side (device, DOF order); x_out: where
    // the solution is copied (device, may be null); x0: initial guess (device,
    // may be null: the problem's resident solution, 0 before its first solve).
    // The caller keeps the buffers alive until wait() returns for the ticket.
    int submit(int64_t pid, const float* b, float* x_out, const float* x0, float tolerance, int64_t* ticket);
    // Blocks until the ticket completes; status is the solver's (0 SUCCESS,
    // 1 FAILED, < 0 error); times in milliseconds from submission.
    // A negative status comes with the failed step and its HIP error in *error.
    int wait(int64_t ticket, int* status, int* iterations, double* queue_ms, double* total_ms,
             int* mg_status = nullptr, int* mg_iterations = nullptr, std::string* error = nullptr);
    ServerStats stats();
    const std::string& error() const { return err_; }

private:
    using clock = std::chrono::steady_clock;
    struct Problem {
        bool visc = false;
        std::unique_ptr<PoissonDevice> pd;
        std::unique_ptr<ViscDevice> vd;
        int64_t ndof = 0;
        size_t bytes = 0;
        bool resident = false;
        bool busy = false;          // a worker holds it (solving, acquiring or being evicted)
        uint64_t last_use = 0;
        float* x = nullptr;         // resident solution (device, ndof floats)
        std::vector<float> x_host;  // solution kept across an eviction
        bool has_solution = false;
    };
    struct Request {
        int64_t ticket, pid;
        const float* b;
        float* x_out;
        const float* x0;
        float tol;
        clock::time_point submitted;
    };
    struct Result {
        bool done = false;
        int status = 0, iterations = 0;
        int mg_status = -1, mg_iterations = -1;   // the pure-multigrid fallback's (-1: not run)
        double queue_ms = 0, total_ms = 0;
        std::string error;                        // the failed step, for a negative status
    };

    void worker();
    int run(Problem& p, const Request& r, int* iterations, int* mg_status, int* mg_iterations, std::string& msg);
    int make_resident(Problem& p, std::string& msg);
    void evict(Problem& p);
    // under the lock: picks victims so that `need` more bytes fit; false if impossible now
    bool plan_room(size_t need, std::vector<Problem*>& victims);

    ServerOptions opt_;
    std::mutex mu_;
    std::condition_variable cv_;        // queue / completion changes
    std::map<int64_t, std::unique_ptr<Problem>> problems_;
    std::deque<Request> queue_;
    std::map<int64_t, Result> results_;
    std::vector<std::thread> workers_;
    int64_t next_pid_ = 1, next_ticket_ = 1;
    uint64_t use_clock_ = 0;
    size_t budget_ = 0, resident_bytes_ = 0;
    int64_t inflight_ = 0;
    ServerStats st_;
    bool stop_ = false;
    std::string err_;
};

}  // namespace uaamg
