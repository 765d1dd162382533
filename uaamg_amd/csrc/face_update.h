// The pressure-gradient update of the face velocities on the device
// (FLIP_Solver_OpenVDB::apply_pressure_gradient, FLIP_vdb.cpp:2571-2667): per
// active face of channel c with weight w > 0 and a pressure DOF on either side,
//   du = (-dt) * (p(v) - p(v - e_c)) / dx / theta,
//   theta = 1, or fraction_inside(phi(v - e_c), phi(v)) clamped to >= 0.02 when
//   either phi >= 0; u += du; reflected = u + du; where w < 1 the friction blend
//   u = (1 - s) u + s u_solid, s = (1 - w) * friction; other active faces are
//   turned off. Operations in the reference's order (bit-identical).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "poisson_kernels.h"

namespace uaamg {

enum { kGradPhi = 0, kGradPhiLow = 1, kGradDof = 2, kGradDofLow = 3, kGradW = 4, kGradSolid = 5, kGradTab = 6 };

struct GradArgs {
    const int32_t* tab;        // [nleaf][kGradTab] leaves of the inputs at the velocity leaf / its -e_c neighbour
    const float* phi;          // [n][512]
    const float* w;            // [n][512][3]
    const float* sv;           // [n][512][3]
    const float* pressure;     // DOF order
    const uint64_t* dmask;     // level-0 DOF masks [nleaf0][8]
    const int64_t* dofbase;    // level-0 first DOF per leaf
    float phi_bg, w_bg[3], sv_bg[3];
    int32_t channel;
    float dt, dx, friction;
};

// active: [nleaf][8] in/out (device); vel in/out; upd / refl may be null
hipError_t launch_pressure_gradient(int64_t nleaf, const GradArgs& a, uint64_t* active, float* vel, float* upd,
                                    float* refl, hipStream_t s);

// A grid as the caller holds it (origins on the host; values host or device).
struct FaceGridRef {
    int64_t nleaf = 0;
    const int32_t* origins = nullptr;
    const float* values = nullptr;
    int32_t ncomp = 1;
    float bg[3] = {0.f, 0.f, 0.f};
};

struct FaceChannelRef {
    int64_t nleaf = 0;
    const int32_t* origins = nullptr;
    uint64_t* active = nullptr;   // host
    float* values = nullptr;
    float* update = nullptr;
    float* reflected = nullptr;
};

class PoissonDevice;

// grids: liquid phi, face weights, solid velocity (nleaf 0: zero); channels: 3.
// Returns 0, or a negative uaamg status with err set.
int apply_pressure_gradient_device(PoissonDevice& d, const float* pressure_dev, const FaceGridRef* grids,
                                   FaceChannelRef* channels, float dt, float friction, bool values_on_device,
                                   std::string& err);

}  // namespace uaamg
