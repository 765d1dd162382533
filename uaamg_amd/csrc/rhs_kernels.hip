// Device build of the pressure Poisson right-hand side (rhs_device.h).
#include "rhs_device.h"
#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

__device__ __forceinline__ float grid_value(const RhsGridDev& g, int32_t leaf, int off, int comp) {
    return leaf >= 0 ? g.values[(size_t(leaf) * 512 + off) * size_t(g.ncomp) + comp] : g.bg[comp];
}

// one wavefront per level-0 DOF leaf, one lane per z-row (the term order of
// BuildPoissonRhs::operator(), simd_vdb_poisson.cpp:43-84)
__global__ __launch_bounds__(kBlockThreads) void poisson_rhs_kernel(LevelView L, RhsArgs a, float* out) {
    const int leaf = wave_leaf();
    if (leaf >= L.nleaf) return;
    const int lane = threadIdx.x & 63;
    const int lx = lane >> 3, ly = lane & 7;
    const uint32_t rm = row_mask(L, leaf, lane);
    if (!rm) return;
    const int32_t* t = a.tab + size_t(leaf) * (kRhsGrids * 4);
    int64_t k = row_dof_base(L, leaf, lane);
    for (int z = 0; z < 8; ++z) {
        if (!((rm >> z) & 1u)) continue;
        float rhs = 0.f;
        bool nonzero_weight = false;
        for (int f = 0; f < 6; ++f) {
            const int ch = f >> 1;
            const bool positive = (f & 1) == 0;
            int X = lx + (positive && ch == 0), Y = ly + (positive && ch == 1), Z = z + (positive && ch == 2);
            int nb = 0;   // 0 own leaf, 1..3 the +x / +y / +z neighbour
            if (X == 8) { nb = 1; X = 0; }
            if (Y == 8) { nb = 2; Y = 0; }
            if (Z == 8) { nb = 3; Z = 0; }
            const int off = (X << 6) | (Y << 3) | Z;
            const float weight = grid_value(a.g[kRhsWeight], t[kRhsWeight * 4 + nb], off, ch);
            if (weight != 0.f) nonzero_weight = true;
            const float v = grid_value(a.g[kRhsVx + ch], t[(kRhsVx + ch) * 4 + nb], off, 0);
            const float sv = grid_value(a.g[kRhsSolid], t[kRhsSolid * 4 + nb], off, ch);
            const float flux = a.invdx * (weight * v + (1.0f - weight) * sv);
            if (positive) rhs -= flux;
            else rhs += flux;
        }
        if (!nonzero_weight) rhs = 0.f;
        out[k++] = rhs;
    }
}

}  // namespace

hipError_t launch_poisson_rhs(const LevelView& L0, const RhsArgs& a, float* rhs_dof, hipStream_t s) {
    if (L0.nleaf <= 0) return hipSuccess;
    poisson_rhs_kernel<<<grid_for(L0.nleaf), kBlockThreads, 0, s>>>(L0, a, rhs_dof);
    return hipGetLastError();
}

}  // namespace uaamg
