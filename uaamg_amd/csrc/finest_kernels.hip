// Device per-voxel stage of the level-0 Poisson build (finest_device.h).
#include "finest_device.h"
#include "fraction_inside.h"
#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

// one wavefront per phi leaf, one lane per z-row
__global__ __launch_bounds__(kBlockThreads) void finest_kernel(int64_t nphi, FinestArgs a) {
    const int leaf = wave_leaf();
    if (leaf >= nphi) return;
    const int lane = threadIdx.x & 63;
    const int lx = lane >> 3, ly = lane & 7;
    const uint32_t am = uint32_t(a.pact[size_t(leaf) * 8 + lx] >> (ly * 8)) & 0xffu;
    const int32_t* nb = a.pnb + size_t(leaf) * 6;
    const int32_t* wt = a.wtab + size_t(leaf) * 4;
    const float t = a.term;
    uint32_t dofs = 0;
    for (int z = 0; z < 8; ++z) {
        const int off = (lx << 6) | (ly << 3) | z;
        const size_t at = size_t(leaf) * 512 + off;
        float diag = 6.f * t, txyz[3] = {-t, -t, -t};
        const float phi_here = a.phi[at];
        if (((am >> z) & 1u) && phi_here < 0.f) {
            float diagonal = 0.f;
            float term_xyz[3] = {0.f, 0.f, 0.f};
            for (int f = 0; f < 6; ++f) {
                const int comp = f >> 1;
                const bool positive = (f & 1) == 0;
                const int lc[3] = {lx, ly, z};
                const int step = comp == 0 ? 64 : (comp == 1 ? 8 : 1);
                // weight: at v + e_comp for the + face, at v for the - face
                float weight;
                if (positive) {
                    const bool in = lc[comp] < 7;
                    const int32_t wl = in ? wt[0] : wt[1 + comp];
                    const int wo = in ? off + step : off - 7 * step;
                    weight = wl >= 0 ? a.w[(size_t(wl) * 512 + wo) * 3 + comp] : a.w_bg[comp];
                } else {
                    weight = wt[0] >= 0 ? a.w[(size_t(wt[0]) * 512 + off) * 3 + comp] : a.w_bg[comp];
                }
                // phi at v +- e_comp
                const bool in = positive ? lc[comp] < 7 : lc[comp] > 0;
                float phi_other;
                if (in) {
                    phi_other = a.phi[size_t(leaf) * 512 + (positive ? off + step : off - step)];
                } else {
                    const int32_t pl = nb[2 * comp + (positive ? 1 : 0)];
                    const int po = positive ? off - 7 * step : off + 7 * step;
                    phi_other = pl >= 0 ? a.phi[size_t(pl) * 512 + po] : a.phi_bg;
                }
                const float term = weight * t;
                if (phi_other < 0) {
                    diagonal += term;
                    if (!positive) term_xyz[comp] = -term;
                } else {
                    float theta = fraction_inside2(phi_here, phi_other);
                    if (theta < 0.02f) theta = 0.02f;
                    diagonal += term / theta;
                }
            }
            if (diagonal != 0.f) {
                dofs |= 1u << z;
                diag = diagonal;
                for (int c = 0; c < 3; ++c) txyz[c] = term_xyz[c];
            }
        }
        a.cd[at] = diag;
        a.cx[at] = txyz[0];
        a.cy[at] = txyz[1];
        a.cz[at] = txyz[2];
    }
    // each lane owns byte ly of its row word
    reinterpret_cast<uint8_t*>(a.dmask + size_t(leaf) * 8 + lx)[ly] = uint8_t(dofs);
}

}  // namespace

hipError_t launch_finest(int64_t nphi, const FinestArgs& a, hipStream_t s) {
    if (nphi <= 0) return hipSuccess;
    if (nphi > int64_t(INT32_MAX) / 64) return hipErrorInvalidValue;
    finest_kernel<<<grid_for(int32_t(nphi)), kBlockThreads, 0, s>>>(nphi, a);
    return hipGetLastError();
}

}  // namespace uaamg
