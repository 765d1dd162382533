// Synthetic free-surface scenes on [0,1]^3 with n^3 cells, flattened to leaf
// bricks in tree order: liquid SDF, staggered face weights, face velocities.
//
// Voxel (i,j,k) is centred at (i,j,k)*dx (the voxel-centre transform of
// FLIP_vdb.cpp:2117-2123). Face weights follow calculate_face_weights
// (FLIP_vdb.cpp:1990-2068): component c of voxel v is the open fraction of the
// face between v-e_c and v, i.e. 1 - fraction_inside of the solid SDF at the
// four face corners, clamped to [0,1]. Faces on or beyond the domain walls have
// weight 0 (solid box walls).
#pragma once

#include <cstdint>
#include <vector>

namespace uaamg {

struct SceneSpec {
    int32_t n = 128;
    float pool = 0.4f;                   // liquid where y < pool (<= 0: no pool)
    std::vector<float> spheres;          // liquid spheres, 4 floats each: cx cy cz r
    bool has_solid = false;
    float solid[4] = {0.5f, 0.5f, 0.5f, 0.2f};
    uint64_t velocity_seed = 0;
    float band_voxels = 2.0f;            // liquid phi narrow band kept as leaves
};

struct Scene {
    SceneSpec spec;
    float dx = 0.f;
    float phi_background = 0.f;
    std::vector<int32_t> origins;   // [nleaf][3]
    std::vector<uint64_t> active;   // [nleaf][8]
    std::vector<float> phi;         // [nleaf][512]
    std::vector<float> weights;     // [nleaf][512][3]
    std::vector<float> vel[3];      // [nleaf][512]
    int64_t nleaf() const { return int64_t(origins.size() / 3); }
};

// counter-based RNG: uniform in [0,1) keyed by (seed, a, b, c, d)
float hash_uniform(uint64_t seed, uint64_t a, uint64_t b, uint64_t c, uint64_t d);

// fraction_inside (levelset_util.cpp:4-103)
float fraction_inside(float phi_left, float phi_right);
float fraction_inside(float phi_bl, float phi_br, float phi_tl, float phi_tr);

void build_scene(const SceneSpec& spec, Scene& out);

// Inputs of the variational viscosity solve (simd_viscosity3d ctor,
// simd_viscosity3d.h:285-290) for a scene: the liquid SDF of build_scene; the
// solid SDF on the vertex lattice (index i at world (i - 1/2) dx, the
// m_voxel_vertex_transform of FLIP_vdb.cpp:2130-2131) = the container walls
// of the domain box (positive inside, FLIP_vdb.cpp:2135-2162) united with the
// solid sphere when the scene has one, kept as leaves where |phi| < 3 dx
// (background 3 dx); a constant viscosity; face velocities U[-1,1] (channel
// c of voxel v at world (v - e_c/2) dx) on the liquid leaves and their 26
// neighbours; zero solid velocity.
struct ViscScene {
    Scene liquid;
    float mu = 1.0f, dt = 1.0f / 24.0f, rho = 1000.0f;
    float solid_background = 0.f;
    std::vector<int32_t> solid_origins;
    std::vector<uint64_t> solid_active;
    std::vector<float> solid_phi;
    std::vector<int32_t> vel_origins;
    std::vector<float> vel;          // [nleaf][512][3]
};

void build_visc_scene(const SceneSpec& spec, float mu, float dt, float rho, ViscScene& out);

// Preset geometries: 0 = pool 40% + sphere r=0.15 at (0.5,0.7,0.5);
// 1 = pool 30% + 8 seeded spheres r in [0.05,0.2] + solid sphere r=0.2 at the
// centre; 3 = one randomised batch problem (edge log-uniform in [32,192],
// 1-6 spheres over a 10-50% pool, solid sphere with probability 1/2).
// n <= 0 keeps the preset's own edge (config 3) or 128.
bool scene_preset(int config, int32_t n, uint64_t seed, SceneSpec& out);

}  // namespace uaamg
