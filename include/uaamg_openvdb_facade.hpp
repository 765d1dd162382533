/*
 * uaamg_openvdb_facade.hpp - the simd_uaamg solver classes over the C ABI (uaamg_capi.h), with the
 * reference's class names, member names and signatures, so that the reference's callers compile
 * against this header unchanged:
 *   - FLIP_Solver_OpenVDB::solve_pressure_simd (FLIP_vdb.cpp:3132-3157): createPressurePoissonLaplacian,
 *     PoissonSolver, mRelativeTolerance, mSmoother, createPressurePoissonRightHandSide,
 *     getZeroVectorGrid, mNumDof, solveMultigridPCG, mIterationTaken, setGridToConstant,
 *     solvePureMultigrid;
 *   - FLIP_Solver_OpenVDB::solve_viscosity (FLIP_vdb.cpp:3178-3208): simd_viscosity3d's grid
 *     constructor, m_matrix_levels[0]->get_zero_vec(), the m_dof_manager[i]->foreach warm-start
 *     visitor, pcg_solve, m_ndof, m_iteration, and packed_FloatGrid3 (v[3], setName, deepCopy, swap,
 *     from_vec3, to_vec3).
 *
 * Header-only; it needs the OpenVDB headers (which this repository's image does not have), so it is
 * compiled only in a build that has them, linked against libuaamg_amd.so. Every solve runs on the
 * GPU through the C ABI; the grids are flattened to leaves on the way in (any leaf order; the library
 * puts them in tree order) and written back leaf by leaf on the way out. The classes mirror the parts
 * of the reference interface its callers use: the per-level operators (laplacianApply, mucycle, ...)
 * are the library's (uaamg_poisson_apply, uaamg_viscosity_apply), not members here, and
 * mMultigridHierarchy / m_matrix_levels hold level 0 only (the coarse levels live in the handle).
 * Errors the reference cannot report (a malformed grid, no device, a HIP failure) throw
 * std::runtime_error with uaamg_last_error()'s text.
 */
#ifndef UAAMG_OPENVDB_FACADE_HPP
#define UAAMG_OPENVDB_FACADE_HPP

#include <openvdb/openvdb.h>
#include <openvdb/tree/LeafManager.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "uaamg_capi.h"

namespace uaamg_facade {

inline void check(int rc, const char* what) {
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + uaamg_last_error());
}

// A grid flattened into the C ABI's uaamg_grid (owns the arrays the struct points to; get()
// points the struct at them, so a moved or copied FlatGrid stays valid).
struct FlatGrid {
    std::vector<int32_t> origins;
    std::vector<uint64_t> active;
    std::vector<float> values;
    uaamg_grid g{};

    void seal(int ncomp, const float* bg) {
        g.nleaf = int64_t(origins.size() / 3);
        g.ncomp = ncomp;
        for (int c = 0; c < 3; ++c) g.background[c] = bg[ncomp == 1 ? 0 : c];
    }
    const uaamg_grid* get() {
        g.origins = origins.data();
        g.active = active.data();
        g.values = values.data();
        return &g;
    }
};

template <class LeafT>
inline void push_leaf_topology(const LeafT& leaf, FlatGrid& f) {
    const openvdb::Coord o = leaf.origin();
    f.origins.push_back(o.x());
    f.origins.push_back(o.y());
    f.origins.push_back(o.z());
    uint64_t words[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (openvdb::Index i = 0; i < LeafT::SIZE; ++i)
        if (leaf.isValueOn(i)) words[i >> 6] |= uint64_t(1) << (i & 63);
    f.active.insert(f.active.end(), words, words + 8);
}

inline FlatGrid flatten(const openvdb::FloatGrid& grid) {
    FlatGrid f;
    for (auto it = grid.tree().cbeginLeaf(); it; ++it) {
        push_leaf_topology(*it, f);
        for (openvdb::Index i = 0; i < 512; ++i) f.values.push_back(it->getValue(i));
    }
    const float bg = grid.background();
    f.seal(1, &bg);
    return f;
}

inline FlatGrid flatten(const openvdb::Vec3fGrid& grid) {
    FlatGrid f;
    for (auto it = grid.tree().cbeginLeaf(); it; ++it) {
        push_leaf_topology(*it, f);
        for (openvdb::Index i = 0; i < 512; ++i) {
            const openvdb::Vec3f v = it->getValue(i);
            f.values.push_back(v[0]);
            f.values.push_back(v[1]);
            f.values.push_back(v[2]);
        }
    }
    const openvdb::Vec3f b = grid.background();
    const float bg[3] = {b[0], b[1], b[2]};
    f.seal(3, bg);
    return f;
}

// leaves of a FloatTree: origins [n][3], active masks [n][8] (bit b of word w = offset 64w+b),
// values [n][512]; every other value is the background
inline openvdb::FloatTree::Ptr make_float_tree(int64_t n, const int32_t* origins, const uint64_t* active,
                                               const float* values, float background) {
    auto tree = std::make_shared<openvdb::FloatTree>(background);
    for (int64_t i = 0; i < n; ++i) {
        auto* leaf = tree->touchLeaf(openvdb::Coord(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]));
        const uint64_t* m = active + 8 * i;
        const float* v = values + 512 * i;
        for (openvdb::Index k = 0; k < 512; ++k) {
            leaf->setValueOnly(k, v[k]);
            if ((m[k >> 6] >> (k & 63)) & 1u) leaf->setValueOn(k);
            else leaf->setValueOff(k);
        }
    }
    return tree;
}

// the DOF topology of level 0 as an Int32 tree holding DOF indices (the reference's mDofIndex /
// m_velocity_DOF.idx[c]); dofbase = -1 leaves are left out
inline openvdb::Int32Tree::Ptr make_dof_tree(int64_t n, const int32_t* origins, const uint64_t* masks,
                                             const int64_t* dofbase, int stride, int channel) {
    auto tree = std::make_shared<openvdb::Int32Tree>(-1);
    for (int64_t i = 0; i < n; ++i) {
        int64_t d = dofbase[i * stride + channel];
        if (d < 0) continue;
        auto* leaf = tree->touchLeaf(openvdb::Coord(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]));
        const uint64_t* m = masks + 8 * (i * stride + channel);
        for (openvdb::Index k = 0; k < 512; ++k) {
            if ((m[k >> 6] >> (k & 63)) & 1u) leaf->setValueOn(k, int32_t(d++));
            else leaf->setValueOff(k, -1);
        }
    }
    return tree;
}

}  // namespace uaamg_facade

// packed_FloatGrid3 (packed3grids.h:5-23): the three face-velocity channels of a staggered grid.
struct packed_FloatGrid3 {
    packed_FloatGrid3() {
        for (int i = 0; i < 3; i++) {
            v[i] = openvdb::FloatGrid::create(0.f);
            v[i]->setName("vc" + std::to_string(i));
        }
        m_transform = openvdb::math::Transform::createLinearTransform(1.0);
        m_gridclass = openvdb::GRID_STAGGERED;
    }
    void setName(std::string name) {
        for (int i = 0; i < 3; i++) v[i]->setName(name + "c" + std::to_string(i));
    }
    // the channel split of uaamg_packed3_from_vec3; a staggered input shifts channel i's transform
    // by -dx/2 along axis i
    void from_vec3(openvdb::Vec3fGrid::Ptr in_v, bool topologycopy = false) {
        m_transform = in_v->transformPtr()->copy();
        uaamg_facade::FlatGrid f = uaamg_facade::flatten(*in_v);
        const int64_t n = f.g.nleaf;
        std::vector<float> ch(size_t(3 * n) * 512);
        if (!topologycopy) uaamg_facade::check(uaamg_packed3_from_vec3(f.get(), ch.data()), "from_vec3");
        for (int i = 0; i < 3; i++) {
            const float bg = in_v->background()[i];
            if (topologycopy) std::fill(ch.begin() + size_t(i * n) * 512, ch.begin() + size_t((i + 1) * n) * 512, bg);
            if (!v[i]) v[i] = openvdb::FloatGrid::create();
            v[i]->setTree(uaamg_facade::make_float_tree(n, f.origins.data(), f.active.data(),
                                                        ch.data() + size_t(i * n) * 512, bg));
            openvdb::math::Transform::Ptr t = in_v->transformPtr()->copy();
            if (in_v->getGridClass() == openvdb::GRID_STAGGERED) {
                openvdb::Vec3d shift(0.0, 0.0, 0.0);
                shift[i] = -0.5 * in_v->voxelSize()[0];
                t->postTranslate(shift);
            }
            v[i]->setTransform(t);
            v[i]->setName(in_v->getName() + "c" + std::to_string(i));
        }
    }
    // the channel merge of uaamg_packed3_to_vec3 into out_v's tree (its transform is kept)
    void to_vec3(openvdb::Vec3fGrid::Ptr out_v) const {
        if (!out_v) return;
        for (int i = 0; i < 3; i++)
            if (!v[i]) return;
        uaamg_facade::FlatGrid f[3] = {uaamg_facade::flatten(*v[0]), uaamg_facade::flatten(*v[1]),
                                       uaamg_facade::flatten(*v[2])};
        int64_t n = 0;
        uaamg_facade::check(uaamg_packed3_to_vec3(f[0].get(), f[1].get(), f[2].get(), &n, nullptr, nullptr, nullptr),
                            "to_vec3");
        std::vector<int32_t> org(static_cast<size_t>(3 * n));
        std::vector<uint64_t> act(static_cast<size_t>(8 * n));
        std::vector<float> val(size_t(n) * 512 * 3);
        uaamg_facade::check(uaamg_packed3_to_vec3(f[0].get(), f[1].get(), f[2].get(), &n, org.data(), act.data(),
                                                  val.data()),
                            "to_vec3");
        auto tree = std::make_shared<openvdb::Vec3fTree>(openvdb::Vec3f(0.f, 0.f, 0.f));
        for (int64_t i = 0; i < n; ++i) {
            auto* leaf = tree->touchLeaf(openvdb::Coord(org[3 * i], org[3 * i + 1], org[3 * i + 2]));
            for (openvdb::Index k = 0; k < 512; ++k) {
                const float* p = &val[(size_t(i) * 512 + k) * 3];
                leaf->setValueOnly(k, openvdb::Vec3f(p[0], p[1], p[2]));
                if ((act[8 * i + (k >> 6)] >> (k & 63)) & 1u) leaf->setValueOn(k);
                else leaf->setValueOff(k);
            }
        }
        out_v->setTree(tree);
    }
    void swap(packed_FloatGrid3& other) {
        for (int i = 0; i < 3; i++) v[i].swap(other.v[i]);
        m_transform.swap(other.m_transform);
    }
    packed_FloatGrid3 deepCopy() const {
        packed_FloatGrid3 result;
        for (int i = 0; i < 3; i++) result.v[i] = v[i]->deepCopy();
        return result;
    }

    openvdb::FloatGrid::Ptr v[3];
    openvdb::GridClass m_gridclass;
    openvdb::math::Transform::Ptr m_transform;
};

namespace simd_uaamg {

// LaplacianWithLevel (simd_vdb_poisson.h:13-131) over a uaamg_poisson handle.
struct LaplacianWithLevel {
    using Ptr = std::shared_ptr<LaplacianWithLevel>;

    static Ptr createPressurePoissonLaplacian(openvdb::FloatGrid::Ptr in_liquid_phi,
                                              openvdb::Vec3fGrid::Ptr in_face_weights, const float in_dt) {
        return std::make_shared<LaplacianWithLevel>(in_liquid_phi, in_face_weights, in_dt,
                                                    float(in_liquid_phi->voxelSize()[0]));
    }

    LaplacianWithLevel(openvdb::FloatGrid::Ptr in_liquid_phi, openvdb::Vec3fGrid::Ptr in_face_weights,
                       const float in_dt, const float in_dx)
        : mDt(in_dt), mDxThisLevel(in_dx), mLevel(0), mTransform(in_liquid_phi->transformPtr()) {
        uaamg_facade::FlatGrid phi = uaamg_facade::flatten(*in_liquid_phi);
        uaamg_facade::FlatGrid w = uaamg_facade::flatten(*in_face_weights);
        uaamg_options o;
        uaamg_facade::check(uaamg_options_default(&o), "options");
        uaamg_poisson* h = nullptr;
        uaamg_facade::check(uaamg_poisson_create(phi.get(), w.get(), in_dt, in_dx, &o, &h),
                            "createPressurePoissonLaplacian");
        mHandle.reset(h, uaamg_poisson_destroy);
        mNumDof = int(uaamg_poisson_ndof(h));
        int64_t nd = 0, nl = 0, nt = 0;
        uaamg_facade::check(uaamg_poisson_level_info(h, 0, &nd, &nl, &nt), "level_info");
        mOrigins.resize(size_t(3 * nl));
        mMasks.resize(size_t(8 * nl));
        mDofBase.resize(size_t(nl));
        uaamg_facade::check(uaamg_poisson_level_topology(h, 0, mOrigins.data(), mMasks.data(), mDofBase.data()),
                            "level_topology");
        mDofIndex = openvdb::Int32Grid::create(uaamg_facade::make_dof_tree(nl, mOrigins.data(), mMasks.data(),
                                                                           mDofBase.data(), 1, 0));
        mDofIndex->setTransform(mTransform->copy());
        mDofLeafManager = std::make_unique<openvdb::tree::LeafManager<openvdb::Int32Tree>>(mDofIndex->tree());
    }

    // uaamg_poisson_build_rhs, written onto the DOF topology
    openvdb::FloatGrid::Ptr createPressurePoissonRightHandSide(openvdb::Vec3fGrid::Ptr in_face_weight,
                                                               openvdb::FloatGrid::Ptr in_vx,
                                                               openvdb::FloatGrid::Ptr in_vy,
                                                               openvdb::FloatGrid::Ptr in_vz,
                                                               openvdb::Vec3fGrid::Ptr in_solid_velocity,
                                                               float /*in_dt: the reference does not use it*/) {
        uaamg_facade::FlatGrid w = uaamg_facade::flatten(*in_face_weight);
        uaamg_facade::FlatGrid vx = uaamg_facade::flatten(*in_vx), vy = uaamg_facade::flatten(*in_vy),
                               vz = uaamg_facade::flatten(*in_vz);
        std::vector<float> rhs(static_cast<size_t>(mNumDof));
        if (in_solid_velocity) {
            uaamg_facade::FlatGrid sv = uaamg_facade::flatten(*in_solid_velocity);
            uaamg_facade::check(uaamg_poisson_build_rhs(mHandle.get(), w.get(), vx.get(), vy.get(), vz.get(),
                                                        sv.get(), rhs.data()),
                                "createPressurePoissonRightHandSide");
        } else {
            uaamg_facade::check(uaamg_poisson_build_rhs(mHandle.get(), w.get(), vx.get(), vy.get(), vz.get(),
                                                        nullptr, rhs.data()),
                                "createPressurePoissonRightHandSide");
        }
        openvdb::FloatGrid::Ptr out = openvdb::FloatGrid::create(0.f);
        vectorToGrid(out, rhs);
        return out;
    }

    // a zero grid on the DOF topology (DOF voxels active)
    openvdb::FloatGrid::Ptr getZeroVectorGrid() {
        const size_t nl = mDofBase.size();
        std::vector<float> zero(nl * 512, 0.f);
        openvdb::FloatGrid::Ptr g = openvdb::FloatGrid::create(0.f);
        g->setTree(uaamg_facade::make_float_tree(int64_t(nl), mOrigins.data(), mMasks.data(), zero.data(), 0.f));
        g->setTransform(mTransform->copy());
        return g;
    }

    // every value of the grid's leaves at the DOF leaves set to constant
    void setGridToConstant(openvdb::FloatGrid::Ptr in_out_grid, float constant) {
        for (size_t i = 0; i < mDofBase.size(); ++i) {
            auto* leaf = in_out_grid->tree().probeLeaf(
                openvdb::Coord(mOrigins[3 * i], mOrigins[3 * i + 1], mOrigins[3 * i + 2]));
            if (leaf) leaf->fill(constant);
        }
    }

    void gridToVector(std::vector<float>& out_vector, openvdb::FloatGrid::Ptr in_grid) {
        out_vector.assign(size_t(mNumDof), 0.f);
        uaamg_facade::FlatGrid g = uaamg_facade::flatten(*in_grid);
        uaamg_facade::check(uaamg_poisson_grid_to_vector(mHandle.get(), g.get(), out_vector.data()), "gridToVector");
    }

    void vectorToGrid(openvdb::FloatGrid::Ptr out_grid, const std::vector<float>& in_vector) {
        std::vector<float> bricks(mDofBase.size() * 512);
        uaamg_facade::check(uaamg_poisson_vector_to_grid(mHandle.get(), in_vector.data(), bricks.data()),
                            "vectorToGrid");
        out_grid->setTree(uaamg_facade::make_float_tree(int64_t(mDofBase.size()), mOrigins.data(), mMasks.data(),
                                                        bricks.data(), 0.f));
        out_grid->setTransform(mTransform->copy());
    }

    // the DOF values of in_vector written into out_grid's leaves (which must hold the DOF topology)
    void vectorToGridUnsafe(openvdb::FloatGrid::Ptr out_grid, const std::vector<float>& in_vector) {
        for (size_t i = 0; i < mDofBase.size(); ++i) {
            auto* leaf = out_grid->tree().probeLeaf(
                openvdb::Coord(mOrigins[3 * i], mOrigins[3 * i + 1], mOrigins[3 * i + 2]));
            if (!leaf) throw std::runtime_error("vectorToGridUnsafe: grid lacks a DOF leaf");
            int64_t d = mDofBase[i];
            const uint64_t* m = &mMasks[8 * i];
            for (openvdb::Index k = 0; k < 512; ++k)
                if ((m[k >> 6] >> (k & 63)) & 1u) leaf->setValueOn(k, in_vector[size_t(d++)]);
        }
    }

    // (row, col, value) triplets of level 0 (getTriplets, simd_vdb_poisson.cpp:129-206); TripletT is
    // constructible from (int, int, float), e.g. Eigen::Triplet<float>
    template <class TripletT>
    void getTriplets(std::vector<TripletT>& out_triplets) const {
        int64_t n = 0;
        uaamg_facade::check(uaamg_poisson_level_triplets(mHandle.get(), 0, &n, nullptr, nullptr, nullptr),
                            "getTriplets");
        std::vector<int32_t> r(static_cast<size_t>(n)), c(static_cast<size_t>(n));
        std::vector<float> v(static_cast<size_t>(n));
        uaamg_facade::check(uaamg_poisson_level_triplets(mHandle.get(), 0, &n, r.data(), c.data(), v.data()),
                            "getTriplets");
        out_triplets.clear();
        out_triplets.reserve(size_t(n));
        for (int64_t k = 0; k < n; ++k) out_triplets.emplace_back(int(r[size_t(k)]), int(c[size_t(k)]), v[size_t(k)]);
    }

    uaamg_poisson* handle() const { return mHandle.get(); }

    openvdb::Int32Grid::Ptr mDofIndex;
    std::unique_ptr<openvdb::tree::LeafManager<openvdb::Int32Tree>> mDofLeafManager;
    int mNumDof = 0;
    float mDt;
    float mDxThisLevel;
    int mLevel;

private:
    openvdb::math::Transform::ConstPtr mTransform;
    std::shared_ptr<uaamg_poisson> mHandle;
    std::vector<int32_t> mOrigins;
    std::vector<uint64_t> mMasks;
    std::vector<int64_t> mDofBase;
};

// PoissonSolver (simd_vdb_poisson.h:133-205): solves on the GPU through the level-0 handle.
class PoissonSolver {
public:
    enum class SmootherOption { ScheduledRelaxedJacobi, WeightedJacobi, RedBlackGaussSeidel, SPAI0 };
    using SuccessType = int;
    static constexpr SuccessType SUCCESS = UAAMG_SUCCESS;
    static constexpr SuccessType FAILED = UAAMG_FAILED;

    PoissonSolver(LaplacianWithLevel::Ptr in_finest_level_matrix) {
        mMultigridHierarchy.push_back(in_finest_level_matrix);
        mIterationTaken = 0;
        mMaxIteration = 100;
        mRelativeTolerance = 1e-7f;
        mSmoother = SmootherOption::ScheduledRelaxedJacobi;
    }

    SuccessType solveMultigridPCG(openvdb::FloatGrid::Ptr in_out_presssure, openvdb::FloatGrid::Ptr in_rhs) {
        return solve(in_out_presssure, in_rhs, true);
    }
    SuccessType solvePureMultigrid(openvdb::FloatGrid::Ptr in_out_presssure, openvdb::FloatGrid::Ptr in_rhs) {
        return solve(in_out_presssure, in_rhs, false);
    }

    int mIterationTaken;
    int mMaxIteration;
    float mRelativeTolerance;
    SmootherOption mSmoother;
    std::vector<LaplacianWithLevel::Ptr> mMultigridHierarchy;

private:
    SuccessType solve(openvdb::FloatGrid::Ptr p, openvdb::FloatGrid::Ptr rhs, bool pcg) {
        LaplacianWithLevel& L = *mMultigridHierarchy[0];
        uaamg_options o;
        uaamg_facade::check(uaamg_options_default(&o), "options");
        o.relative_tolerance = mRelativeTolerance;
        o.max_iterations = mMaxIteration;
        o.smoother = int32_t(mSmoother);
        uaamg_facade::check(uaamg_poisson_set_options(L.handle(), &o), "set_options");
        std::vector<float> x, b;
        L.gridToVector(x, p);
        L.gridToVector(b, rhs);
        int32_t it = 0;
        const int rc = pcg ? uaamg_poisson_solve_pcg(L.handle(), x.data(), b.data(), &it)
                           : uaamg_poisson_solve_mg(L.handle(), x.data(), b.data(), &it);
        uaamg_facade::check(rc, pcg ? "solveMultigridPCG" : "solvePureMultigrid");
        mIterationTaken = it;
        L.vectorToGridUnsafe(p, x);
        return rc;
    }
};

// L_with_level (simd_viscosity3d.h:12-281), level 0, over a uaamg_viscosity handle.
struct L_with_level {
    enum class working_mode { JACOBI, SPAI0, NORMAL, RESIDUAL, RED_GS, BLACK_GS };
    struct DOF_tuple_t {
        openvdb::Int32Grid::Ptr idx[3];
    };
    using Ptr = std::shared_ptr<L_with_level>;

    static Ptr create_viscosity_matrix(openvdb::FloatGrid::Ptr in_viscosity, openvdb::FloatGrid::Ptr in_liquid_sdf,
                                       openvdb::FloatGrid::Ptr in_solid_sdf, float in_dt, float in_rho) {
        return std::make_shared<L_with_level>(in_viscosity, in_liquid_sdf, in_solid_sdf, in_dt, in_rho);
    }

    L_with_level(openvdb::FloatGrid::Ptr in_viscosity, openvdb::FloatGrid::Ptr in_liquid_sdf,
                 openvdb::FloatGrid::Ptr in_solid_sdf, float in_dt, float in_rho)
        : m_dx(float(in_liquid_sdf->voxelSize()[0])), m_dt(in_dt), m_rho(in_rho), m_level(0),
          m_viscosity(in_viscosity), m_solid_sdf(in_solid_sdf), m_liquid_sdf(in_liquid_sdf) {
        uaamg_facade::FlatGrid mu = uaamg_facade::flatten(*in_viscosity);
        uaamg_facade::FlatGrid liq = uaamg_facade::flatten(*in_liquid_sdf);
        uaamg_facade::FlatGrid sol = uaamg_facade::flatten(*in_solid_sdf);
        uaamg_viscosity_options o;
        uaamg_facade::check(uaamg_viscosity_options_default(&o), "options");
        uaamg_viscosity* h = nullptr;
        uaamg_facade::check(uaamg_viscosity_create(mu.get(), liq.get(), sol.get(), in_dt, in_rho, m_dx, &o, &h),
                            "create_viscosity_matrix");
        m_handle.reset(h, uaamg_viscosity_destroy);
        m_ndof = int(uaamg_viscosity_ndof(h));
        int64_t nd = 0, nchan[3] = {0, 0, 0}, nl = 0;
        uaamg_facade::check(uaamg_viscosity_level_info(h, 0, &nd, nchan, &nl), "level_info");
        m_origins.resize(size_t(3 * nl));
        m_masks.resize(size_t(24 * nl));
        m_dofbase.resize(size_t(3 * nl));
        uaamg_facade::check(uaamg_viscosity_level_topology(h, 0, m_origins.data(), m_masks.data(), m_dofbase.data()),
                            "level_topology");
        m_voxel_center_transform = in_liquid_sdf->transformPtr()->copy();
        for (int c = 0; c < 3; ++c) {
            m_face_center_transform[c] = m_voxel_center_transform->copy();
            openvdb::Vec3d shift(0.0, 0.0, 0.0);
            shift[c] = -0.5 * double(m_dx);
            m_face_center_transform[c]->postTranslate(shift);
            m_velocity_DOF.idx[c] = openvdb::Int32Grid::create(
                uaamg_facade::make_dof_tree(nl, m_origins.data(), m_masks.data(), m_dofbase.data(), 3, c));
            m_velocity_DOF.idx[c]->setTransform(m_face_center_transform[c]);
            m_dof_manager[c] =
                std::make_unique<openvdb::tree::LeafManager<openvdb::Int32Tree>>(m_velocity_DOF.idx[c]->tree());
        }
    }

    // uaamg_viscosity_build_rhs_packed, written onto the DOF topology
    packed_FloatGrid3 build_rhs(packed_FloatGrid3 in_velocity_field, openvdb::Vec3fGrid::Ptr in_solid_velocity) const {
        return to_packed_FloatGrid3(rhs_vector(in_velocity_field, in_solid_velocity));
    }

    std::vector<float> rhs_vector(const packed_FloatGrid3& u, openvdb::Vec3fGrid::Ptr in_solid_velocity) const {
        uaamg_facade::FlatGrid v0 = uaamg_facade::flatten(*u.v[0]), v1 = uaamg_facade::flatten(*u.v[1]),
                               v2 = uaamg_facade::flatten(*u.v[2]);
        std::vector<float> rhs(static_cast<size_t>(m_ndof));
        if (in_solid_velocity) {
            uaamg_facade::FlatGrid sv = uaamg_facade::flatten(*in_solid_velocity);
            uaamg_facade::check(uaamg_viscosity_build_rhs_packed(m_handle.get(), v0.get(), v1.get(), v2.get(),
                                                                 sv.get(), rhs.data()),
                                "build_rhs");
        } else {
            uaamg_facade::check(uaamg_viscosity_build_rhs_packed(m_handle.get(), v0.get(), v1.get(), v2.get(),
                                                                 nullptr, rhs.data()),
                                "build_rhs");
        }
        return rhs;
    }

    // to_Eigenvector (simd_viscosity3d.cpp:2089-2117) into any vector with setZero(n) and operator[]
    // (Eigen::VectorXf), or into a std::vector<float>
    void to_vector(std::vector<float>& out, const packed_FloatGrid3& in) const {
        out.assign(size_t(m_ndof), 0.f);
        uaamg_facade::FlatGrid v0 = uaamg_facade::flatten(*in.v[0]), v1 = uaamg_facade::flatten(*in.v[1]),
                               v2 = uaamg_facade::flatten(*in.v[2]);
        uaamg_facade::check(uaamg_viscosity_grid_to_vector(m_handle.get(), v0.get(), v1.get(), v2.get(), out.data()),
                            "to_Eigenvector");
    }
    template <class VecT>
    void to_Eigenvector(VecT& out_eigenvector, packed_FloatGrid3 in_float3grid) const {
        std::vector<float> v;
        to_vector(v, in_float3grid);
        out_eigenvector.setZero(m_ndof);
        for (int d = 0; d < m_ndof; ++d) out_eigenvector[d] = v[size_t(d)];
    }

    packed_FloatGrid3 to_packed_FloatGrid3(const std::vector<float>& in_vector) const {
        const size_t nl = m_dofbase.size() / 3;
        std::vector<float> bricks(nl * 3 * 512);
        uaamg_facade::check(uaamg_viscosity_vector_to_grid(m_handle.get(), in_vector.data(), bricks.data()),
                            "write_to_FloatGrid3");
        packed_FloatGrid3 result;
        std::vector<float> vals(nl * 512);
        std::vector<uint64_t> masks(nl * 8);
        for (int c = 0; c < 3; ++c) {
            for (size_t i = 0; i < nl; ++i) {
                std::copy(&bricks[(i * 3 + size_t(c)) * 512], &bricks[(i * 3 + size_t(c)) * 512] + 512, &vals[i * 512]);
                std::copy(&m_masks[(i * 3 + size_t(c)) * 8], &m_masks[(i * 3 + size_t(c)) * 8] + 8, &masks[i * 8]);
            }
            result.v[c]->setTree(uaamg_facade::make_float_tree(int64_t(nl), m_origins.data(), masks.data(),
                                                               vals.data(), 0.f));
            result.v[c]->setTransform(m_face_center_transform[c]->copy());
            result.v[c]->setName("vc" + std::to_string(c));
        }
        return result;
    }

    // the DOF values written into out_grids' leaves, which hold the DOF topology
    // (write_to_FloatGrid3_assume_topology, simd_viscosity3d.cpp:2141-2154)
    void write_to_FloatGrid3_assume_topology(packed_FloatGrid3 out_grids, const std::vector<float>& in_vector) const {
        const size_t nl = m_dofbase.size() / 3;
        for (int c = 0; c < 3; ++c)
            for (size_t i = 0; i < nl; ++i) {
                int64_t d = m_dofbase[i * 3 + size_t(c)];
                if (d < 0) continue;
                auto* leaf = out_grids.v[c]->tree().probeLeaf(
                    openvdb::Coord(m_origins[3 * i], m_origins[3 * i + 1], m_origins[3 * i + 2]));
                if (!leaf) throw std::runtime_error("write_to_FloatGrid3_assume_topology: grid lacks a DOF leaf");
                const uint64_t* m = &m_masks[(i * 3 + size_t(c)) * 8];
                for (openvdb::Index k = 0; k < 512; ++k)
                    if ((m[k >> 6] >> (k & 63)) & 1u) leaf->setValueOn(k, in_vector[size_t(d++)]);
            }
    }

    packed_FloatGrid3 get_zero_vec() const { return to_packed_FloatGrid3(std::vector<float>(size_t(m_ndof), 0.f)); }

    uaamg_viscosity* handle() const { return m_handle.get(); }

    float m_dx;
    float m_dt;
    float m_rho;
    int m_level;
    openvdb::FloatGrid::Ptr m_viscosity;
    openvdb::FloatGrid::Ptr m_solid_sdf;
    openvdb::FloatGrid::Ptr m_liquid_sdf;
    DOF_tuple_t m_velocity_DOF;
    int m_ndof = 0;
    std::unique_ptr<openvdb::tree::LeafManager<openvdb::Int32Tree>> m_dof_manager[3];
    openvdb::math::Transform::Ptr m_voxel_center_transform;
    openvdb::math::Transform::Ptr m_face_center_transform[3];

private:
    std::shared_ptr<uaamg_viscosity> m_handle;
    std::vector<int32_t> m_origins;
    std::vector<uint64_t> m_masks;     // [nleaf][3][8]
    std::vector<int64_t> m_dofbase;    // [nleaf][3]
};

// simd_viscosity3d (simd_viscosity3d.h:283-320): pcg_solve on the GPU through the level-0 handle.
struct simd_viscosity3d {
    simd_viscosity3d(openvdb::FloatGrid::Ptr in_viscosity, openvdb::FloatGrid::Ptr in_liquid_sdf,
                     openvdb::FloatGrid::Ptr in_solid_sdf, packed_FloatGrid3 in_liquid_velocity,
                     openvdb::Vec3fGrid::Ptr in_solid_velocity, float in_dt, float in_rho)
        : simd_viscosity3d(L_with_level::create_viscosity_matrix(in_viscosity, in_liquid_sdf, in_solid_sdf, in_dt,
                                                                 in_rho),
                           in_liquid_velocity, in_solid_velocity) {}

    simd_viscosity3d(L_with_level::Ptr level0, packed_FloatGrid3 in_liquid_velocity,
                     openvdb::Vec3fGrid::Ptr in_solid_velocity) {
        m_matrix_levels.push_back(level0);
        m_rhs_vector = level0->rhs_vector(in_liquid_velocity, in_solid_velocity);
        m_rhs = level0->to_packed_FloatGrid3(m_rhs_vector);
        m_iteration = 0;
        m_max_iter = 100;
    }

    // in_lhs: the initial guess on the DOF topology (get_zero_vec's), overwritten with the result
    void pcg_solve(packed_FloatGrid3 in_lhs, float tolerance) {
        L_with_level& L = *m_matrix_levels[0];
        uaamg_viscosity_options o;
        uaamg_facade::check(uaamg_viscosity_options_default(&o), "options");
        o.max_iterations = m_max_iter;
        uaamg_facade::check(uaamg_viscosity_set_options(L.handle(), &o), "set_options");
        std::vector<float> x;
        L.to_vector(x, in_lhs);
        int32_t it = 0;
        uaamg_facade::check(uaamg_viscosity_solve_pcg(L.handle(), tolerance, x.data(), m_rhs_vector.data(), &it),
                            "pcg_solve");
        m_iteration = it;
        L.write_to_FloatGrid3_assume_topology(in_lhs, x);
    }

    std::vector<std::shared_ptr<L_with_level>> m_matrix_levels;
    packed_FloatGrid3 m_rhs;
    int m_iteration;
    int m_max_iter;

private:
    std::vector<float> m_rhs_vector;
};

}  // namespace simd_uaamg

#endif  // UAAMG_OPENVDB_FACADE_HPP
