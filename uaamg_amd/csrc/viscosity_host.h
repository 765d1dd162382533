// Host construction of the variational-viscosity level hierarchy in leaf-brick
// form (simd_viscosity3d.cpp).
//
// Level 0 follows L_with_level's finest constructor (:22-50): subcell SDF and
// liquid volumes (:258-418), DOF marking (:420-537), the 15-point
// coefficients (:539-908). Coarse levels follow L_with_level(child,
// Coarsening) (:52-78, 1017-1451). The hierarchy loop and the finest-level
// trimming follow the simd_viscosity3d constructor (:3621-3658): coarsen while
// ndof > 4000, then trim level 0 (eps 1e-3).
//
// One leaf table per level holds the three velocity channels: a vector is
// [leaf][channel][512] floats, leaves holding a DOF of any channel first (in
// tree order), then leaves that only hold coefficients. Each coefficient grid
// of the reference is a brick per leaf in a shared pool, or slot -1 where the
// reference has no leaf (or trimmed it): the level operator then reads the
// level's default constant.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "leaf_topology.h"

namespace uaamg {

// coefficient grid ids: diag[c], invdiag[c], neg[c][axis] (coupling along
// -axis of channel c), cross[normal][j][k] (m_cross_channel), SPAI0[c]
enum ViscGrid { kVDiag = 0, kVInvDiag = 3, kVNeg = 6, kVCross = 15, kVSpai0 = 27, kVNumGrids = 30, kVSlots = 32 };
inline int vneg(int c, int axis) { return kVNeg + 3 * c + axis; }
inline int vcross(int normal, int j, int k) { return kVCross + 4 * normal + 2 * j + k; }

struct ViscLevelHost {
    int level = 0;
    float dx = 0.f, dt = 0.f, rho = 0.f;
    float default_diag = 1.f, default_off = 1.f;
    float dflt[kVSlots] = {};        // level operator's constant for a missing leaf
    int64_t ndof = 0;
    int64_t nchan[3] = {0, 0, 0};

    std::vector<Coord> origin;       // DOF leaves first, then coefficient-only leaves
    size_t nleaf_dof = 0;
    LeafLookup lookup;
    std::vector<uint64_t> mask;      // [nleaf_dof][3][8]
    std::vector<int64_t> dofbase;    // [nleaf_dof][3] first DOF index of the channel in the leaf (-1 none)
    std::vector<int32_t> nb27;       // [nleaf][27] leaf at offset (ax,ay,az) in {-1,0,1}^3, -1 none
    std::vector<int32_t> slot;       // [nleaf][kVSlots] brick in pool or -1
    BrickVec pool;                   // [nslot][512]
    std::vector<uint64_t> pool_on;   // [nslot][8] positions the reference keeps active
    std::vector<int32_t> chan_leaves[3];   // leaves holding DOFs of channel c
    std::vector<int32_t> parent;     // [nleaf_dof] coarse leaf * 8 + octant (not on the coarsest level)
    std::vector<int32_t> child;      // [nleaf_dof][8] finer DOF leaf per octant (not on level 0)
    bool spai0 = false;

    size_t nleaf() const { return origin.size(); }
    // value the level operator reads for grid gid at voxel off of leaf i
    float coef(int gid, size_t leaf, int off) const {
        const int32_t s = slot[leaf * kVSlots + gid];
        return s < 0 ? dflt[gid] : pool[size_t(s) * kLeafSize + off];
    }
    bool is_dof(size_t leaf, int c, int off) const {
        return leaf < nleaf_dof && mask_test(&mask[(leaf * 3 + c) * 8], off);
    }
};

// finest-level geometry kept for the right-hand side (build_rhs, :1724-2087)
struct ViscGeometry {
    float dx = 0.f, dt = 0.f, rho = 0.f;
    // voxel / face / edge liquid volumes on the liquid leaves dilated twice
    std::vector<Coord> vol_origin;
    LeafLookup vol_lookup;
    std::vector<float> voxel_vol, face_vol[3], edge_vol[3];   // [n][512]
    // solid faces and DOF indices on the liquid leaves dilated once (6 neighbours)
    std::vector<Coord> dof_origin;
    LeafLookup dof_lookup;
    std::vector<uint64_t> solid_on[3];   // [n][8]
    std::vector<int32_t> dof_index[3];   // [n][512], -1 none
    HostGrid mu;
};

struct ViscHierarchyHost {
    std::vector<ViscLevelHost> levels;
    ViscGeometry geo;
    // explicit matrix of the coarsest level (CSR, columns ascending)
    std::vector<int64_t> cptr;
    std::vector<int32_t> ccol;
    std::vector<float> cval;
};

// Builds every level. trim = false keeps level 0 untrimmed (inspection).
// Returns false on bad input.
// volume_device >= 0 runs the level-0 volume stage (subcell SDF, subcell and voxel / face / edge volumes)
// on that HIP device, bit-identical to the host stage (-1); err receives the reason of a failure.
bool build_visc_hierarchy(const HostGrid& liquid, const HostGrid& solid, const HostGrid& mu, float dt, float rho,
                          float dx, int64_t max_coarsest_dof, bool trim, ViscHierarchyHost& out,
                          int volume_device = -1, std::string* err = nullptr);

// SPAI0 of one level (build_SPAI0_matrix / SPAI0_collector, :1453-1698).
void build_visc_spai0(ViscLevelHost& L);

// Rows of a level's explicit matrix (explicit_matrix_reducer, :2415-2596),
// columns ascending as Eigen's setFromTriplets stores them.
void visc_explicit_matrix(const ViscLevelHost& L, std::vector<int64_t>& ptr, std::vector<int32_t>& col,
                          std::vector<float>& val);

// Right-hand side as a DOF-ordered vector (build_rhs, :1942-2087). u and
// solid_u are Vec3f grids with face-centred components; solid_u may be null.
void build_visc_rhs(const ViscHierarchyHost& h, const HostGrid& u, const HostGrid* solid_u, float* out_dof);
// the same with the liquid velocity as three single-channel grids (packed_FloatGrid3's v[0..2])
void build_visc_rhs(const ViscHierarchyHost& h, const HostGrid* const channels[3], const HostGrid* solid_u,
                    float* out_dof);

// volume_fraction of a cube from its 8 corner values (volume_fractions.cpp:100-126)
float visc_volume_fraction(float p000, float p100, float p010, float p110, float p001, float p101, float p011,
                           float p111);

}  // namespace uaamg
