// Device-resident UAAMG-PCG Poisson solver: owns the level hierarchy in HBM and
// drives the W-cycle preconditioner and the outer PCG through the kernels of
// poisson_kernels.hip. Mirrors PoissonSolver (simd_vdb_poisson.h:139-205).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "poisson_host.h"
#include "poisson_kernels.h"
#include "stream_order.h"

namespace uaamg {

enum SmootherOption { kScheduledRelaxedJacobi = 0, kWeightedJacobi = 1, kRedBlackGaussSeidel = 2, kSpai0 = 3 };

constexpr int kSweeps = 3;   // n = 3 pre/post sweeps (solveMultigridPCG, simd_vdb_poisson.cpp:2181,2212)
// Jacobi weight of sweep i of the preconditioner's / the pure multigrid cycle's schedule
float scheduled_weight(int n, int i);
float scheduled_weight_iterative(int n, int i);

// The coarsest level as explicit rows for the single-workgroup CG: per DOF the
// brick position (leaf_offset + leaf) * 512 + offset, the six neighbour DOFs
// and the stencil values (diag, c+x, c-x, c+y, c-y, c+z, c-z, Eigen's invdiag).
void coarsest_system(const PoissonLevelHost& l, int64_t leaf_offset, std::vector<int32_t>& pos,
                     std::vector<int32_t>& nb, std::vector<float>& coef);

struct SolverOptions {
    float relative_tolerance = 1e-7f;
    int max_iterations = 100;
    int smoother = kScheduledRelaxedJacobi;
    int precision = 0;   // 0: FP32 throughout; 1: BF16-stored W-cycle vectors + FP64 residual replacement
    int device = 0;
    int64_t max_coarsest_dof = 4000;
    int flags = 0;   // bit 0: unfused kernels, bit 1: no graph capture, bit 2: tail kernel (A-B checks),
                     // bit 4: no block-visit kernels, bit 6: FP64 residual replacement at FP32
    hipStream_t stream = nullptr;
};

// One device allocation carved into aligned pieces.
class DeviceArena {
public:
    ~DeviceArena();
    hipError_t reserve(size_t bytes);
    void release();
    void* take(size_t bytes);
    size_t used() const { return used_; }
    size_t capacity() const { return cap_; }

private:
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
};

class PoissonDevice {
public:
    static constexpr int kSuccess = 0;
    static constexpr int kFailed = 1;
    static constexpr int kMaxReplacements = 4;
    static constexpr int kTuneMaxLeaves = 4096;   // coarse levels above this size keep the per-sweep kernels

    ~PoissonDevice();
    // returns 0 or a negative uaamg status
    int init(PoissonHierarchyHost&& h, const SolverOptions& opt);
    // keeps the host hierarchy only (inspection without a GPU); solves fail
    void init_host_only(PoissonHierarchyHost&& h, const SolverOptions& opt);
    bool on_device() const { return !lv_.empty(); }
    // device residency: the bytes acquire_device() reserves (one arena, sized
    // from the levels' leaf counts), uploading the kept host hierarchy, and
    // freeing every device resource while keeping the host hierarchy
    size_t device_bytes() const;
    int acquire_device();
    // keep_host: first copy device-held coefficient bricks to the host, so a
    // later acquire_device() can upload them again
    void release_device(bool keep_host = true);

    int nlevels() const { return int(host_.levels.size()); }
    const PoissonHierarchyHost& host() const { return host_; }
    SolverOptions& options() { return opt_; }
    hipStream_t stream() const { return stream_; }
    // orders the solver's stream after the caller's null-stream work (own stream only)
    int follow_caller();
    int device() const { return opt_.device; }
    // device memory for the steps either side of the solve (leaf tables, masks): grown on demand and
    // kept until the device resources are released, so repeated calls allocate nothing; the caller
    // has made the handle's device current. Returns null when the allocation fails.
    void* aux(size_t bytes);
    // device view of a level (on_device() only)
    const LevelView& level_view(int l) const { return lv_[size_t(l)].v; }
    const std::string& error() const { return err_; }

    // Solves with DOF-ordered vectors in device memory. x is the initial guess
    // on entry and the solution on exit. Returns kSuccess/kFailed or < 0.
    int solve_pcg_dev(float* x_dof, const float* b_dof, int* iterations);
    int solve_mg_dev(float* x_dof, const float* b_dof, int* iterations);

    // unit operations on DOF-ordered device vectors (parity tests)
    int apply_dev(int level, int mode, float w, const float* x_dof, const float* b_dof, float* y_dof);
    int first_sweep_dev(int level, int kind, float w, const float* b_dof, float* y_dof);
    int restrict_dev(int fine_level, const float* fine_dof, float* coarse_dof);
    int prolong_dev(int fine_level, const float* coarse_dof, float* fine_dof, float alpha);
    int precondition_dev(const float* r_dof, float* z_dof);
    int coarsest_dev(const float* b_dof, float* x_dof);

    // scratch DOF-ordered device vectors of level-0 size (host API staging)
    float* staging(int i) { return staging_[i]; }

    // Bytes one sweep of `mode` on `level` must move at minimum: every stream
    // read or written once per DOF, coefficient streams only for leaves that
    // keep stored coefficients (trimmed leaves read constants).
    double algorithmic_bytes(int level, int mode) const;
    // Mean duration of `reps` back-to-back launches of one operator sweep on
    // the solver's stream, timed with HIP events on that stream.
    int time_apply(int level, int mode, int reps, float* ms_per_launch);

    // Level-to-kernel mapping chosen at acquire time: per level, whether a
    // W-cycle visit runs as two block-visit launches, and the two timings the
    // choice came from (ms per visit pair; -1 where the level was not timed).
    const std::vector<uint8_t>& block_levels() const { return use_block_; }
    // the block post-visit of level l also runs the coarsest solve
    bool post_cg_level(int l) const {
        return post_cg_ && l >= 1 && l + 2 == nlevels() && size_t(l) < use_block_.size() && use_block_[size_t(l)] &&
               l + 1 != tail_level_;
    }
    float tune_ms(int level, int variant) const {
        return level < int(tune_ms_[variant].size()) ? tune_ms_[variant][size_t(level)] : -1.f;
    }

    // ||b - A x||_inf of the last solve's result (FP64 residual) when residual
    // replacement is on, and how many replacements it made; -1 otherwise
    float last_true_residual() const { return true_nu_; }
    // SPAI0 coefficients, built (host) and uploaded (device) on first use
    int ensure_spai0();
    // Copies every level's coefficient bricks to the host hierarchy where they
    // live only on the device (from the build's device store, or from the
    // packed arrays of the resident hierarchy); 0 or negative with error() set.
    int ensure_host_coefficients();
    int last_replacements() const { return replacements_; }

private:
    struct DevLevel {
        LevelView v{};
        float* X = nullptr;
        float* B = nullptr;
        float* T = nullptr;
        float* U = nullptr;   // second smoothing buffer of the fused W-cycle
        // the same vectors in BF16 (mixed-precision form only)
        bf16* Xh = nullptr;
        bf16* Bh = nullptr;
        bf16* Th = nullptr;
        bf16* Uh = nullptr;
    };

    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
    };
    int capture(Graph& g, const std::function<void()>& body);
    void issue_init();
    void issue_first();
    void issue_iter(int k, bool replace = false);
    int ensure_graphs();
    void destroy_graphs();
    int run(Graph& g, const std::function<void()>& body);
    int wait_nu(float* nu);

    int upload_level(const PoissonLevelHost& h, DevLevel& d, bool spai0);
    std::vector<int32_t> level_parents(int l) const;
    int build_coarsest();
    void precond(int level, float* X, const float* B, bool skip_first, int dot_target);
    void precond_fused(int level, float* X, const float* B, bool skip_first, int dot_target);
    template <class TX, class TB>
    void precond_mixed(int level, TX* X, const TB* B, bool skip_first, int dot_target);
    int pcg_iterate(float& nu, float numax, int& it, bool& nu_true, float restart_below, bool replace);
    void tune_levels();
    void visit_pair(int level, bool block);
    void mg_cycle(int level, float* X, const float* B);
    void smooth(int level, float*& cur, float*& oth, const float* B, int sweep, bool pre, bool iterative);
    int check(hipError_t e, const char* what);

    PoissonHierarchyHost host_;
    SolverOptions opt_;
    hipStream_t stream_ = nullptr;
    CallerOrder caller_;
    bool own_stream_ = false;
    void* aux_ = nullptr;
    size_t aux_bytes_ = 0;
    DeviceArena arena_;
    std::vector<DevLevel> lv_;
    std::vector<float> sor_w_;
    // level-0 PCG bricks
    float *x0_ = nullptr, *b0_ = nullptr, *r0_ = nullptr, *z0_ = nullptr, *q0_ = nullptr;
    float* p_[2] = {nullptr, nullptr};   // search direction, ping-pong across iterations
    float* staging_[3] = {nullptr, nullptr, nullptr};
    PcgScalars* sc_ = nullptr;
    PcgScalars* sc_host_ = nullptr;   // pinned mirror
    double* partials_ = nullptr;
    CoarsestView cv_{};
    const int32_t* coarsest_dof_of_ = nullptr;   // [coarsest leaves * 512]: DOF at each brick position or -1
    bool post_cg_ = false;                 // block post-visit of level L-2 runs the coarsest solve in its launch
    // captured PCG pieces, valid for graph_smoother_ / graph_sor_
    Graph g_init_, g_first_, g_iter_[2];
    Graph g_repl_[2];   // g_iter_ with the FP64 residual replacement (captured with refine_)
    int graph_smoother_ = -1;
    std::vector<void*> spai0_bufs_;   // SPAI0 bricks uploaded after acquire_device
    bool spai0_on_device_ = false;
    std::vector<float> graph_sor_;
    bool use_graphs_ = true;
    bool fuse_ = true;                     // fused W-cycle / PCG kernels
    int tail_level_ = -1;                  // level whose visit pairs run in the tail kernel (-1: none)
    int block_max_leaves_ = 0;             // threshold mode: coarse levels up to this size use the block-visit kernels
    std::vector<uint8_t> use_block_;       // per level: visits run as two block-visit launches
    std::vector<float> tune_ms_[2];        // per level: measured ms of a visit pair, [0] block-visit, [1] per-sweep
    bool mixed_ = false;                   // BF16-stored W-cycle vectors (Jacobi-family smoothers)
    bool refine_ = false;                  // FP64 residual replacement
    bool graph_refine_ = false;
    float true_nu_ = -1.f;
    int replacements_ = -1;
    HostMailbox* mailbox_ = nullptr;       // pinned host memory
    HostMailbox* mailbox_dev_ = nullptr;   // its device alias
    unsigned int nu_seq_ = 0;              // ||r||_inf reductions issued so far
    hipError_t first_error_ = hipSuccess;
    std::string err_;
};

}  // namespace uaamg
