// Batched variational-viscosity solve: many independent problems of uneven
// size merged into one leaf table per level, so every colour pass, transfer
// and PCG vector operation is one launch for the whole batch and one hipGraph
// per PCG iteration covers all problems (the viscosity form of PoissonBatch).
//
// Merged layout. Problems are ordered by level count, deepest first. Level l
// holds the DOF leaves of every problem that has the level, in that order
// (level 0 starts every problem at a multiple of the reduction block and pads
// it to one), then their coefficient-only leaves. Leaf indices in the
// neighbourhood, parent and child tables and in the wave records are offset
// into the merged numbering; the coefficient pools are concatenated, and every
// slot that reads the level default (an absent leaf or brick) points instead
// at a brick of the problem's own default value, so problems with different
// dt, rho, dx or viscosity background share one launch. The problems that
// smooth at level l (it is not their coarsest) are a prefix: their colour
// passes and residual run on a view with their wave records only; problems
// whose coarsest level is l solve it there, one workgroup each. Level-0
// reductions are per problem in the single-problem block order, so each
// problem's iterate is bit-identical to its solve alone
// (simd_viscosity3d::pcg_solve, simd_viscosity3d.cpp:3694-3740).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "batch_kernels.h"
#include "poisson_kernels.h"
#include "poisson_solver.h"
#include "stream_order.h"
#include "viscosity_host.h"
#include "viscosity_kernels.h"
#include "viscosity_solver.h"

namespace uaamg {

class ViscBatch {
public:
    ~ViscBatch();
    // Merges and uploads the problems (FP32 form). opt: max_iterations,
    // device, flags (bit 1: no graph replay), stream. 0 or a negative uaamg
    // status.
    int init(std::vector<ViscHierarchyHost>&& problems, const ViscOptions& opt);

    int nproblems() const { return int(order_.size()); }
    int64_t ndof(int input_index) const;
    int nlevels() const { return int(lv_.size()); }
    int64_t level_leaves(int level) const { return level < nlevels() ? lv_[size_t(level)].v.nleaf : 0; }
    size_t device_bytes() const { return arena_.capacity(); }
    const std::string& error() const { return err_; }
    ViscOptions& options() { return opt_; }

    // pcg_solve(tolerance) on every problem: x[i] / b[i] are DOF-ordered
    // device vectors of input problem i (x: initial guess in, result out);
    // iterations[i] = its m_iteration. Returns 0 or a negative uaamg status.
    int solve(float tolerance, float* const* x, const float* const* b, int* iterations);

private:
    struct Level {
        ViscLevelView v{};    // every problem present at this level
        ViscLevelView sv{};   // the problems that smooth here (wave records of the prefix only)
        LevelView flat{};     // 3 * nleaf_dof bricks (level 0: the PCG vector kernels)
        float *X = nullptr, *B = nullptr, *T = nullptr;
        int coarse_first = 0, coarse_count = 0;   // problems whose coarsest level this is (merged order)
        int coarse_max_n = 0;
    };
    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
    };

    int check(hipError_t e, const char* what);
    void rbgs(int l, float* x, const float* b, bool xyz);
    void mucycle(int l, float* X, const float* B, bool first);
    void issue_init();
    void issue_first();
    void issue_iter();
    void issue_final();
    int capture(Graph& g, const std::function<void()>& body);
    int run(Graph& g, const std::function<void()>& body);
    int wait_active(int* active);
    void destroy_graphs();

    ViscOptions opt_;
    hipStream_t stream_ = nullptr;
    CallerOrder caller_;
    bool own_stream_ = false;
    DeviceArena arena_;
    std::vector<Level> lv_;
    std::vector<int> order_;              // merged position -> input index
    std::vector<int> rank_;               // input index -> merged position
    std::vector<int> nlev_;               // level count per merged problem
    std::vector<int64_t> ndof_;           // level-0 DOFs per merged problem
    std::vector<LevelView> sub0_;         // level-0 flat view of each merged problem alone (scatter / gather)
    std::vector<int64_t> off0_;           // level-0 brick offset (floats) of each merged problem
    const int32_t* brickprob_ = nullptr;  // level-0 flat brick -> merged problem
    const int32_t* blk_begin_ = nullptr;  // [nprob + 1] first reduction block per merged problem
    ViscCoarsestView* views_ = nullptr;   // device array, merged order
    BatchPcg* st_ = nullptr;
    std::vector<BatchPcg> st_host_;
    unsigned int* seq_dev_ = nullptr;
    double* partials_ = nullptr;
    float *x0_ = nullptr, *b0_ = nullptr, *r0_ = nullptr, *p0_ = nullptr, *z0_ = nullptr;
    Graph g_init_, g_first_, g_iter_, g_final_;
    bool graphs_ready_ = false;
    bool use_graphs_ = true;
    HostMailbox* mailbox_ = nullptr;
    HostMailbox* mailbox_dev_ = nullptr;
    unsigned int seq_ = 0;
    hipError_t first_error_ = hipSuccess;
    std::string err_;
};

}  // namespace uaamg
