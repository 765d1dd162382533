// Device pressure-gradient update of the face velocities (face_update.h).
#include "face_update.h"
#include "fraction_inside.h"
#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

// pressure at a voxel of level-0 DOF leaf dl: (is a DOF, value; 0 off the DOFs)
__device__ __forceinline__ bool pressure_at(const GradArgs& a, int32_t dl, int off, float& p) {
    p = 0.f;
    if (dl < 0) return false;
    const uint64_t* m = a.dmask + size_t(dl) * 8;
    const int w = off >> 6, b = off & 63;
    if (!((m[w] >> b) & 1u)) return false;
    int64_t k = a.dofbase[dl];
    for (int i = 0; i < w; ++i) k += __popcll(m[i]);
    k += __popcll(b ? (m[w] & ((uint64_t(1) << b) - 1)) : 0ull);
    p = a.pressure[k];
    return true;
}

// one wavefront per velocity leaf, one lane per z-row
__global__ __launch_bounds__(kBlockThreads) void pressure_gradient_kernel(int64_t nleaf, GradArgs a, uint64_t* active,
                                                                          float* vel, float* upd, float* refl) {
    const int leaf = wave_leaf();
    if (leaf >= nleaf) return;
    const int lane = threadIdx.x & 63;
    const int lx = lane >> 3, ly = lane & 7;
    const int c = a.channel;
    const int32_t* t = a.tab + size_t(leaf) * kGradTab;
    const uint32_t rm = uint32_t(active[size_t(leaf) * 8 + lx] >> (ly * 8)) & 0xffu;
    uint32_t out_mask = 0;
    for (int z = 0; z < 8; ++z) {
        const int off = (lx << 6) | (ly << 3) | z;
        const size_t at = size_t(leaf) * 512 + off;
        const float u0 = vel[at];
        float updated = u0, reflected = u0, du = 0.f;
        bool has_update = false;
        if ((rm >> z) & 1u) {
            const float weight = t[kGradW] >= 0 ? a.w[(size_t(t[kGradW]) * 512 + off) * 3 + c] : a.w_bg[c];
            if (weight > 0.0f) {
                // the voxel below along c: same leaf, or the -e_c neighbour at local coordinate 7
                const int lc = c == 0 ? lx : (c == 1 ? ly : z);
                const bool below_in = lc > 0;
                const int step = c == 0 ? 64 : (c == 1 ? 8 : 1);
                const int off_b = below_in ? off - step : off + 7 * step;
                float p_this, p_below;
                const bool has_p = pressure_at(a, t[kGradDof], off, p_this);
                const bool has_pb = pressure_at(a, below_in ? t[kGradDof] : t[kGradDofLow], off_b, p_below);
                if (has_p || has_pb) {
                    has_update = true;
                    const int32_t pl = below_in ? t[kGradPhi] : t[kGradPhiLow];
                    const float phi_this = t[kGradPhi] >= 0 ? a.phi[size_t(t[kGradPhi]) * 512 + off] : a.phi_bg;
                    const float phi_below = pl >= 0 ? a.phi[size_t(pl) * 512 + off_b] : a.phi_bg;
                    float theta = 1.0f;
                    if (phi_this >= 0 || phi_below >= 0) {
                        theta = fraction_inside2(phi_below, phi_this);
                        if (theta < 0.02f) theta = 0.02f;
                    }
                    du = -a.dt * (p_this - p_below) / a.dx / theta;
                    updated += du;
                    reflected = updated + du;
                    if (weight < 1.0f) {
                        const float svel =
                            t[kGradSolid] >= 0 ? a.sv[(size_t(t[kGradSolid]) * 512 + off) * 3 + c] : a.sv_bg[c];
                        const float solid_fraction = (1.0f - weight) * a.friction;
                        updated = (1.0f - solid_fraction) * updated + solid_fraction * svel;
                    }
                }
            }
        }
        if (has_update) {
            out_mask |= 1u << z;
            vel[at] = updated;
        }
        if (upd) upd[at] = has_update ? du : u0;
        if (refl) refl[at] = has_update ? reflected : u0;
    }
    // each lane owns byte ly of its row word
    reinterpret_cast<uint8_t*>(active + size_t(leaf) * 8 + lx)[ly] = uint8_t(out_mask);
}

}  // namespace

hipError_t launch_pressure_gradient(int64_t nleaf, const GradArgs& a, uint64_t* active, float* vel, float* upd,
                                    float* refl, hipStream_t s) {
    if (nleaf <= 0) return hipSuccess;
    if (nleaf > int64_t(INT32_MAX) / 64) return hipErrorInvalidValue;
    pressure_gradient_kernel<<<grid_for(int32_t(nleaf)), kBlockThreads, 0, s>>>(nleaf, a, active, vel, upd, refl);
    return hipGetLastError();
}

}  // namespace uaamg
