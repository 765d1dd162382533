// Batched UAAMG-PCG Poisson solve: many independent problems of uneven size
// in one merged leaf table per level, so every W-cycle stage and every PCG
// vector operation is one launch for the whole batch, and one hipGraph per
// PCG iteration covers all problems.
//
// Merged layout. Problems are ordered by level count, deepest first. Level l
// of the merged table holds, in that order, level l of every problem that has
// one; the problems that smooth at level l (it is not their coarsest) form a
// prefix, so the smoothing kernels run on a prefix view and the transfers on
// the whole level. Leaf indices in the neighbour, 27-neighbourhood and child
// tables are offset into the merged numbering (no table crosses problems).
// Coefficients: each problem's kept bricks plus one brick per coefficient kind
// holding its default value, which the slots of its trimmed leaves point to
// (the per-problem form of "absent reads the default"). Level 0 starts every
// problem at a multiple of the reduction block and pads it to one, so the
// reductions, the per-problem finalize and hence every iterate equal a solve
// of that problem alone. Problems whose coarsest level is l get their coarsest
// CG there, one workgroup each. Each problem keeps its own PCG scalars and
// state on the device; a problem that converges or fails stops updating while
// the others continue, and the host waits on one mailbox word that counts the
// problems still iterating.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "batch_kernels.h"
#include "poisson_host.h"
#include "poisson_kernels.h"
#include "poisson_solver.h"

namespace uaamg {

class PoissonBatch {
public:
    ~PoissonBatch();
    // Merges and uploads the problems. opt: tolerance, max_iterations,
    // smoother (Jacobi family only), precision, device, flags (bit 1: no graph
    // replay, bit 4: no block-visit kernels, bit 7: a problem whose PCG fails
    // is solved again from zero by pure multigrid, as FLIP does). 0 or a
    // negative uaamg status.
    int init(std::vector<PoissonHierarchyHost>&& problems, const SolverOptions& opt);

    int nproblems() const { return int(order_.size()); }
    int64_t ndof(int input_index) const;
    int nlevels() const { return int(lv_.size()); }
    int64_t level_leaves(int level) const { return level < nlevels() ? lv_[level].v.nleaf : 0; }
    size_t device_bytes() const { return arena_.capacity(); }
    const std::string& error() const { return err_; }
    SolverOptions& options() { return opt_; }

    // x[i] / b[i]: DOF-ordered device vectors of input problem i (x: initial
    // guess in, solution out). status[i] = 0 SUCCESS / 1 FAILED, iterations[i]
    // = its mIterationTaken. Returns 0 or a negative uaamg status.
    int solve(float* const* x, const float* const* b, int* status, int* iterations);
    // after a solve with residual replacement (BF16 W-cycle, or flags bit 6):
    // input problem i's FP64 true residual and restart count (-1 without)
    float last_true_residual(int i) const { return i < int(last_.size()) ? last_[size_t(i)].true_nu : -1.f; }
    int last_replacements(int i) const { return i < int(last_.size()) ? last_[size_t(i)].replacements : -1; }
    // after a solve with the pure-multigrid fallback (flags bit 7): input
    // problem i's solvePureMultigrid status and iterations, -1 where its PCG
    // did not fail (or the fallback is off)
    int last_mg_status(int i) const { return i < int(mg_status_.size()) ? mg_status_[size_t(i)] : -1; }
    int last_mg_iterations(int i) const { return i < int(mg_iters_.size()) ? mg_iters_[size_t(i)] : -1; }

private:
    struct Level {
        LevelView v{};             // every problem present at this level
        LevelView sv{};            // prefix of the problems that smooth here
        float *X = nullptr, *B = nullptr, *T = nullptr, *U = nullptr;
        bf16 *Xh = nullptr, *Bh = nullptr, *Th = nullptr, *Uh = nullptr;
        int coarse_first = 0, coarse_count = 0;   // problems whose coarsest level this is (merged order)
        int coarse_max_ndof = 0;                  // largest of their coarsest systems
    };
    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
    };

    int check(hipError_t e, const char* what);
    void precond(int l, float* X, const float* B, bool first_visit, int dot_op);
    template <class TX, class TB>
    void precond_mixed(int l, TX* X, const TB* B, bool first_visit, int dot_op);
    void level0_dot(float* X, const float* B, int dot_op);
    void issue_init();
    void issue_first();
    void issue_iter(int k);
    int capture(Graph& g, const std::function<void()>& body);
    int run(Graph& g, const std::function<void()>& body);
    int wait_active(int* active);
    void destroy_graphs();
    void tune_levels();
    void visit_pair(int level, bool block);
    int fallback(int i, float* x, const float* b);

    // pure-multigrid fallback (flags bit 7): each problem's host hierarchy,
    // kept at init, and a single-problem solver made on its first failure
    std::vector<PoissonHierarchyHost> fb_host_;
    std::vector<std::unique_ptr<PoissonDevice>> fb_dev_;
    std::vector<int> mg_status_, mg_iters_;

    SolverOptions opt_;
    hipStream_t stream_ = nullptr;
    CallerOrder caller_;
    bool own_stream_ = false;
    DeviceArena arena_;
    std::vector<Level> lv_;
    std::vector<int> order_;              // merged position -> input index
    std::vector<int> rank_;               // input index -> merged position
    std::vector<int> nlev_;               // level count per merged problem
    std::vector<int64_t> ndof_;           // level-0 DOFs per merged problem
    std::vector<int64_t> off0_;           // level-0 leaf offset per merged problem
    std::vector<LevelView> sub0_;         // level-0 view of each merged problem alone (DOF scatter / gather)
    LevelView c0_{};                      // level-0 view of the problems whose only level is level 0
    int64_t c0_block_ = 0;                // first reduction block of that view
    const int32_t* leafprob_ = nullptr;   // level-0 leaf -> merged problem
    const int32_t* blk_begin_ = nullptr;  // [nprob + 1] first reduction block per merged problem
    CoarsestView* views_ = nullptr;       // device array, merged order
    BatchPcg* st_ = nullptr;
    std::vector<BatchPcg> st_host_;
    std::vector<BatchPcg> last_;         // input order, after the last solve
    unsigned int* seq_dev_ = nullptr;
    double* partials_ = nullptr;
    float *x0_ = nullptr, *b0_ = nullptr, *r0_ = nullptr, *z0_ = nullptr, *q0_ = nullptr;
    float* p_[2] = {nullptr, nullptr};
    Graph g_init_, g_first_, g_iter_[2];
    bool graphs_ready_ = false;
    bool use_graphs_ = true;
    bool mixed_ = false;
    int64_t block_max_leaves_ = 0;
    std::vector<uint8_t> use_block_;      // per merged level: visits as two block-visit launches
    HostMailbox* mailbox_ = nullptr;
    HostMailbox* mailbox_dev_ = nullptr;
    unsigned int seq_ = 0;
    hipError_t first_error_ = hipSuccess;
    std::string err_;
};

}  // namespace uaamg
