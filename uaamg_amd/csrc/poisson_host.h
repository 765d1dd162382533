// Host construction of the UAAMG Poisson level hierarchy in leaf-brick form.
//
// Level 0 follows LaplacianWithLevel's finest constructor
// (simd_vdb_poisson.cpp:208-253, BuildFinestMatrix :474-597, initializeFinest
// :600-635, setDofIndex :637-671, trimDefaultNodes :1756-1811). Coarse levels
// follow initializeFromFineLevel (:260-471) with TouchCoarseLeafReducer
// (SIMD_UAAMG_Ops.h:4-37). The hierarchy loop follows
// constructMultigridHierarchy (:1813-1842): coarsen while ndof > 4000.
#pragma once

#include <cstdint>
#include <string>
#include <memory>
#include <vector>

#include "leaf_topology.h"

namespace uaamg {

// Coefficient grids of one level. Values are the effective values the reference
// reads through its accessors after trimming: where a leaf was trimmed, every
// voxel reads the default stencil value.
enum CoefKind { kCoefDiag = 0, kCoefX = 1, kCoefY = 2, kCoefZ = 3, kCoefSpai0 = 4, kNumCoef = 5 };

struct PoissonLevelHost {
    int level = 0;
    float dt = 0.f, dx = 0.f, term = 0.f;   // term = dt / dx^2 at this level

    // DOF topology. Only leaves holding at least one DOF are kept ("compact");
    // empty leaves of the reference tree carry no DOF and no stencil work.
    std::vector<Coord> origin;
    std::vector<uint64_t> mask;      // [nleaf][8]
    std::vector<int64_t> dofbase;    // [nleaf] first DOF index in the leaf
    int64_t ndof = 0;
    LeafLookup lookup;
    std::vector<int32_t> nbr;        // [nleaf][6] face neighbours -x,+x,-y,+y,-z,+z (-1 absent)

    // Reference-tree bookkeeping: the reference keeps empty leaves in its DOF
    // tree; this is that leaf count (level 0: the liquid-phi leaf count).
    std::vector<Coord> full_origin;

    // effective coefficient bricks [nleaf][512] and whether the reference keeps
    // that coefficient leaf (1) or trimmed it to the default (0)
    BrickVec coef[kNumCoef];
    std::vector<uint8_t> keep[kNumCoef];
    BrickVec invdiag;                // [nleaf][512], 1/diag (0 where diag == 0); default where diag trimmed

    // set when the bricks (coef[0..3], invdiag) live only in the hierarchy's
    // device store (build_hierarchy_device); the host arrays are then empty
    bool coef_on_device = false;

    // aggregation links to the finer level (empty on level 0):
    // child[nleaf][8]: finer-level leaf index of octant (bx<<2|by<<1|bz), -1 absent
    std::vector<int32_t> child;

    size_t nleaf() const { return origin.size(); }
    float default_coef(int kind) const;
    float default_invdiag() const { return 1.0f / (6.f * term); }
    // Value the apply operator substitutes for a trimmed SPAI0 leaf
    // (mDefaultSPAI0, simd_vdb_poisson.cpp:1069).
    float default_spai0_apply() const { return 1.0f / (7.f * term); }
};

// Device copies of a hierarchy's coefficient bricks (full [nleaf][512]
// arrays laid out as the host arrays, trimmed leaves holding the default),
// one entry per level; levels kept on the host have null pointers. The
// memory is released with the last owner.
struct DeviceCoefStore {
    struct Level {
        float* coef[4] = {nullptr, nullptr, nullptr, nullptr};   // diag, x, y, z
        float* invdiag = nullptr;
    };
    int device = -1;
    std::vector<Level> levels;
    std::vector<void*> blocks;   // allocations owned by the store
    DeviceCoefStore() = default;
    DeviceCoefStore(const DeviceCoefStore&) = delete;
    DeviceCoefStore& operator=(const DeviceCoefStore&) = delete;
    ~DeviceCoefStore();
};

struct PoissonHierarchyHost {
    std::vector<PoissonLevelHost> levels;
    int64_t nleaf_phi = 0;
    bool spai0_ready = false;
    std::shared_ptr<DeviceCoefStore> dcoef;   // levels with coef_on_device
};

// Level 0 from the liquid phi and face-weight grids. device >= 0 runs the
// per-voxel stage on that HIP device (bit-identical; finest_device.h). Returns
// false on bad input or a device error (err, when given, says which).
bool build_finest_level(const HostGrid& phi, const HostGrid& weights, float dt, float dx, PoissonLevelHost& out,
                        int device = -1, std::string* err = nullptr);

// Next coarser level from a finer one.
void build_coarse_level(const PoissonLevelHost& fine, PoissonLevelHost& out);
// Its topology only: origins (TouchCoarseLeafReducer), child table, DOF masks
// (a coarse voxel is a DOF iff one of its 2^3 fine voxels is), numbering and
// face neighbours; no coefficients.
void build_coarse_topology(const PoissonLevelHost& fine, PoissonLevelHost& out);
// DOF numbering, leaf lookup and face-neighbour table of a level whose
// origins and masks are set.
void finalize_topology(PoissonLevelHost& l);

// Full hierarchy: coarsen while ndof > max_coarsest_dof (device: as build_finest_level).
bool build_hierarchy(const HostGrid& phi, const HostGrid& weights, float dt, float dx, int64_t max_coarsest_dof,
                     PoissonHierarchyHost& out, int device = -1, std::string* err = nullptr);

// SPAI0 coefficients of every level (init_SPAI0, simd_vdb_poisson.cpp:1113-1190).
void build_spai0(PoissonHierarchyHost& h);

// Right-hand side of the pressure Poisson equation as a DOF-ordered vector
// (BuildPoissonRhs, simd_vdb_poisson.cpp:19-93).
void build_poisson_rhs(const PoissonLevelHost& l0, const HostGrid& weights, const HostGrid& vx, const HostGrid& vy,
                       const HostGrid& vz, const HostGrid& solid_vel, float* out_rhs);

// The explicit matrix of a level as getTriplets emits it
// (LaplacianTripletsReducer, simd_vdb_poisson.cpp:129-196): DOFs in order, and
// per DOF (row, row, diag), then per axis x, y, z the lower neighbour DOF with
// the coefficient of the DOF's own face and the upper neighbour DOF with the
// coefficient of the neighbour's face, each only where that neighbour is a DOF.
void level_triplets(const PoissonLevelHost& l, std::vector<int32_t>& rows, std::vector<int32_t>& cols,
                    std::vector<float>& vals);

// Run fn(begin, end) over [0, n) on the host worker threads, at least `grain`
// items per thread (the default suits light per-item work; per-leaf loops over
// 512 voxels pass 1, so that even a few leaves spread over the threads).
template <class Fn>
void parallel_for(int64_t n, Fn&& fn, int64_t grain = 32);

int host_threads();

}  // namespace uaamg

#include <algorithm>
#include <thread>

namespace uaamg {
template <class Fn>
void parallel_for(int64_t n, Fn&& fn, int64_t grain) {
    int nt = host_threads();
    grain = std::max<int64_t>(grain, 1);
    if (n < 2 * grain || nt <= 1) {
        fn(int64_t(0), n);
        return;
    }
    nt = int(std::min<int64_t>(nt, (n + grain - 1) / grain));
    std::vector<std::thread> th;
    th.reserve(nt);
    int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&fn, b, e]() { fn(b, e); });
    }
    for (auto& t : th) t.join();
}
}  // namespace uaamg
