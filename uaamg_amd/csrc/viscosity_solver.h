// Device-resident variational viscosity solver: owns the level hierarchy in
// HBM and drives the W-cycle preconditioner and the outer PCG of
// simd_viscosity3d (simd_viscosity3d.h:283-321, .cpp:3621-3884).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "poisson_kernels.h"
#include "poisson_solver.h"
#include "stream_order.h"
#include "viscosity_host.h"
#include "viscosity_kernels.h"

namespace uaamg {

struct ViscOptions {
    float tolerance = 1e-6f;        // pcg_solve's tolerance argument
    int max_iterations = 100;       // m_max_iter
    int device = 0;
    int64_t max_coarsest_dof = 4000;
    int flags = 0;                  // bit 1: no graph capture, bit 3: untrimmed level 0, bit 5: single-workgroup
                                    // half-visits up to 16 leaves, bit 7: no single-workgroup half-visits
    int precision = 0;              // 0: FP32; 1: the W-cycle's coarse-level vectors stored in BF16
    hipStream_t stream = nullptr;
};

// Wave records of a level (layout at kVRecLeaf), chan_leaves order: per
// (channel, DOF leaf) the leaf, its 27-leaf neighbourhood and the coefficient
// slots its 15-point row reads (-1: absent, the level default).
std::vector<int32_t> visc_wave_records(const ViscLevelHost& h);

// Host arrays of the coarsest level's explicit system (ViscCoarsestView):
// Eigen's diagonal preconditioner, the rows as padded ELL, and each DOF's
// position in the level vector, level leaf i standing at leaf_map[i] (null:
// at i). False with err set when the system exceeds the kernel's limits.
struct ViscCoarsestHost {
    int64_t n = 0;
    int32_t width = 0;
    std::vector<float> inv, eval;
    std::vector<int32_t> pos, ecol, elen;
};
bool visc_coarsest_host(const ViscHierarchyHost& h, const int32_t* leaf_map, ViscCoarsestHost& out,
                        std::string& err);

class ViscDevice {
public:
    ~ViscDevice();
    // returns 0 or a negative uaamg status
    int init(ViscHierarchyHost&& h, const ViscOptions& opt);
    void init_host_only(ViscHierarchyHost&& h, const ViscOptions& opt);
    bool on_device() const { return !lv_.empty(); }
    // device residency (see PoissonDevice): arena size from the levels' leaf
    // counts, upload of the kept host hierarchy, release of every device resource
    size_t device_bytes() const;
    int acquire_device();
    void release_device();

    int nlevels() const { return int(host_.levels.size()); }
    const ViscHierarchyHost& host() const { return host_; }
    ViscHierarchyHost& host_mut() { return host_; }
    ViscOptions& options() { return opt_; }
    hipStream_t stream() const { return stream_; }
    const std::string& error() const { return err_; }

    // pcg_solve on DOF-ordered device vectors: x holds the initial guess on
    // entry and the result on exit; *iterations receives m_iteration.
    int solve_pcg_dev(float* x_dof, const float* b_dof, int* iterations);

    // unit operations on DOF-ordered device vectors (parity tests). Modes
    // 0..3 apply to every channel; 4 / 5 run one colour pass on channels 0,
    // 1, 2 in turn (in place on a copy of x); 6 / 7 are XYZ_RBGS / ZYX_RBGS.
    int apply_dev(int level, int mode, const float* x_dof, const float* b_dof, float* y_dof);
    int restrict_dev(int fine_level, const float* fine_dof, float* coarse_dof);
    int prolong_dev(int fine_level, const float* coarse_dof, float* fine_dof);   // fine += coarse
    int coarsest_dev(const float* b_dof, float* x_dof);
    int precondition_dev(const float* r_dof, float* z_dof);

    float* staging(int i) { return staging_[i]; }
    // orders the solver's stream after the caller's null-stream work (own stream only)
    int follow_caller();

    // level-to-kernel mapping chosen at acquire time (1 = single-workgroup
    // half-visits) and the timings behind it (-1 where not timed)
    const std::vector<uint8_t>& small_levels() const { return use_small_; }
    float tune_ms(int level, int variant) const {
        return level < int(tune_ms_[variant].size()) ? tune_ms_[variant][size_t(level)] : -1.f;
    }
    static constexpr int64_t kSmallTuneLeaves = 64;

    // bytes one pass of `mode` over channel c of `level` must move (each
    // stream once per DOF, coefficient streams only where stored)
    double algorithmic_bytes(int level, int mode, int channel) const;
    // mean ms per launch of `reps` back-to-back passes, HIP events on the stream
    int time_apply(int level, int mode, int channel, int reps, float* ms_per_launch);

private:
    struct DevLevel {
        ViscLevelView v{};
        LevelView flat{};          // the vector as 3*nleaf_dof bricks for the PCG kernels
        // the W-cycle's level vectors: FP32, or BF16 (bf = true) on the coarse
        // levels of the mixed form
        void *X = nullptr, *B = nullptr, *T = nullptr;
        bool bf = false;
        float* Xf() const { return static_cast<float*>(X); }
        float* Bf() const { return static_cast<float*>(B); }
        float* Tf() const { return static_cast<float*>(T); }
    };
    struct Graph {
        hipGraph_t g = nullptr;
        hipGraphExec_t e = nullptr;
    };

    int upload_level(const ViscLevelHost& h, DevLevel& d);
    int upload_spai0(int level);
    int build_coarsest();
    template <class V>
    void rbgs(int level, V* x, const V* b, bool xyz);
    template <class V>
    void mucycle(int level, V* x, const V* b, bool precon);
    template <class V>
    void apply_all(int level, int mode, const V* x, const V* b, V* y);
    template <class V>
    void half_visit(int level, bool small);
    void issue_init();
    void issue_first();
    void issue_iter();
    int capture(Graph& g, const std::function<void()>& body);
    int ensure_graphs();
    void destroy_graphs();
    int run(Graph& g, const std::function<void()>& body);
    int wait_nu(float* nu);
    void tune_levels();
    void half_visit_pair(int level, bool small);
    bool bf_level(int level) const { return opt_.precision == 1 && level >= 1; }
    int check(hipError_t e, const char* what);

    ViscHierarchyHost host_;
    ViscOptions opt_;
    hipStream_t stream_ = nullptr;
    CallerOrder caller_;
    bool own_stream_ = false;
    DeviceArena arena_;
    std::vector<DevLevel> lv_;
    float *x0_ = nullptr, *b0_ = nullptr, *r0_ = nullptr, *p0_ = nullptr, *z0_ = nullptr;
    float* staging_[3] = {nullptr, nullptr, nullptr};
    PcgScalars* sc_ = nullptr;
    double* partials_ = nullptr;
    ViscCoarsestView cv_{};
    std::vector<float*> spai0_bufs_;
    Graph g_init_, g_first_, g_iter_;
    bool graphs_ready_ = false;
    bool use_graphs_ = true;
    int64_t small_max_leaves_ = 0;   // candidate coarse levels: widest channel up to this many leaves
    std::vector<uint8_t> use_small_;  // per level: half-visits as one single-workgroup launch
    std::vector<float> tune_ms_[2];   // per level: ms per half-visit pair, [0] single-workgroup, [1] per-pass
    HostMailbox* mailbox_ = nullptr;
    HostMailbox* mailbox_dev_ = nullptr;
    unsigned int nu_seq_ = 0;
    hipError_t first_error_ = hipSuccess;
    std::string err_;
};

}  // namespace uaamg
