// Device kernels of the viscosity level-0 volume stage (visc_volumes.h): one
// workgroup per leaf, its neighbourhood staged in LDS, every value computed with
// the host stage's operations in the host stage's order.
#include <hip/hip_runtime.h>

#include "visc_volumes_kernels.h"
#include "volume_fraction.h"

namespace uaamg {
namespace {

constexpr int kT = 256;
constexpr int kSubN = 16;                  // subcells per leaf edge
constexpr int kSubCells = kSubN * kSubN * kSubN;

// BoxSampler's lerp: a + T((b - a) * w) with the weight in double
__device__ __forceinline__ float lerp_d(float a, float b, double t) { return a + float(double(b - a) * t); }

// subcell SDF of one leaf: the liquid SDF at voxels o-1 .. o+8 staged, then each
// subcell s = 2o + l sampled at voxel position (s - 1) / 2, clamped to bg
__global__ __launch_bounds__(kT) void subcell_sdf_kernel(VolumeStageArgs a, float* sdf) {
    __shared__ float blk[10 * 10 * 10];
    const int leaf = blockIdx.x;
    const int32_t* nb = a.liq_nb + size_t(leaf) * 27;
    for (int t = threadIdx.x; t < 1000; t += kT) {
        const int i = t / 100, j = (t / 10) % 10, k = t % 10;
        const int X = i - 1, Y = j - 1, Z = k - 1;
        const int ax = X < 0 ? -1 : (X > 7 ? 1 : 0), ay = Y < 0 ? -1 : (Y > 7 ? 1 : 0),
                  az = Z < 0 ? -1 : (Z > 7 ? 1 : 0);
        const int32_t l = nb[(ax + 1) * 9 + (ay + 1) * 3 + (az + 1)];
        const int off = ((X - 8 * ax) << 6) | ((Y - 8 * ay) << 3) | (Z - 8 * az);
        blk[t] = l >= 0 ? a.liquid[size_t(l) * 512 + size_t(off)] : a.liquid_bg;
    }
    __syncthreads();
    float* out = sdf + size_t(leaf) * kSubCells;
    for (int t = threadIdx.x; t < kSubCells; t += kT) {
        const int lx = t >> 8, ly = (t >> 4) & 15, lz = t & 15;
        const int ix = (lx + 1) >> 1, iy = (ly + 1) >> 1, iz = (lz + 1) >> 1;
        const double u = (lx & 1) ? 0.0 : 0.5, v = (ly & 1) ? 0.0 : 0.5, w = (lz & 1) ? 0.0 : 0.5;
        auto B = [&](int x, int y, int z) { return blk[(x * 10 + y) * 10 + z]; };
        const float smp = lerp_d(lerp_d(lerp_d(B(ix, iy, iz), B(ix, iy, iz + 1), w),
                                        lerp_d(B(ix, iy + 1, iz), B(ix, iy + 1, iz + 1), w), v),
                                 lerp_d(lerp_d(B(ix + 1, iy, iz), B(ix + 1, iy, iz + 1), w),
                                        lerp_d(B(ix + 1, iy + 1, iz), B(ix + 1, iy + 1, iz + 1), w), v),
                                 u);
        out[t] = smp < a.sdf_bg ? smp : a.sdf_bg;
    }
}

// subcell liquid fractions of one leaf from the 17^3 corner samples (its own 16^3
// and the +x / +y / +z neighbours' first layers; a missing leaf reads bg)
__global__ __launch_bounds__(kT) void subcell_volume_kernel(VolumeStageArgs a, const float* sdf, float* svol) {
    __shared__ float c17[17 * 17 * 17];
    const int leaf = blockIdx.x;
    const int32_t* nb = a.sub_nb_up + size_t(leaf) * 8;
    for (int t = threadIdx.x; t < 17 * 17 * 17; t += kT) {
        const int i = t / 289, j = (t / 17) % 17, k = t % 17;
        const int32_t l = nb[((i >> 4) << 2) | ((j >> 4) << 1) | (k >> 4)];
        c17[t] = l >= 0 ? sdf[size_t(l) * kSubCells + size_t((((i & 15) << 4) | (j & 15)) << 4 | (k & 15))]
                        : a.sdf_bg;
    }
    __syncthreads();
    float* out = svol + size_t(leaf) * kSubCells;
    for (int t = threadIdx.x; t < kSubCells; t += kT) {
        const int i = t >> 8, j = (t >> 4) & 15, k = t & 15;
        auto P = [&](int x, int y, int z) { return c17[((i + x) * 17 + j + y) * 17 + k + z]; };
        const float f = volume_fraction8(P(0, 0, 0), P(1, 0, 0), P(0, 1, 0), P(1, 1, 0), P(0, 0, 1), P(1, 0, 1),
                                         P(0, 1, 1), P(1, 1, 1));
        out[t] = f > 0 ? f : 0.f;
    }
}

// voxel / face / edge volumes of one leaf: subcells 2o-1 .. 2o+16 staged; each
// centre sums its eight subcells in the host stage's order, times 1/8
__global__ __launch_bounds__(kT) void centre_volume_kernel(VolumeStageArgs a, const float* svol, float* out) {
    __shared__ float s18[18 * 18 * 18];
    const int leaf = blockIdx.x;
    const int32_t* nb = a.vol_sub_nb + size_t(leaf) * 27;
    for (int t = threadIdx.x; t < 18 * 18 * 18; t += kT) {
        const int i = t / 324, j = (t / 18) % 18, k = t % 18;
        // subcell 2o - 1 + i: leaf offset -1 (i = 0), 0 (1..16) or +1 (17), local index (i + 15) & 15
        const int ax = i == 0 ? -1 : (i == 17 ? 1 : 0), ay = j == 0 ? -1 : (j == 17 ? 1 : 0),
                  az = k == 0 ? -1 : (k == 17 ? 1 : 0);
        const int32_t l = nb[(ax + 1) * 9 + (ay + 1) * 3 + (az + 1)];
        const int lx = (i + 15) & 15, ly = (j + 15) & 15, lz = (k + 15) & 15;
        s18[t] = l >= 0 ? svol[size_t(l) * kSubCells + size_t(((lx << 4) | ly) << 4 | lz)] : 0.f;
    }
    __syncthreads();
    const size_t n = size_t(gridDim.x) * 512;
    for (int off = threadIdx.x; off < 512; off += kT) {
        const int vx = off >> 6, vy = (off >> 3) & 7, vz = off & 7;
        auto S = [&](int x, int y, int z) { return s18[((x + 1) * 18 + y + 1) * 18 + z + 1]; };
        float vc = 0, vf[3] = {0, 0, 0}, ve[3] = {0, 0, 0};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) {
                    const int p[3] = {2 * vx + i, 2 * vy + j, 2 * vz + k};
                    vc += S(p[0], p[1], p[2]);
                    for (int c = 0; c < 3; ++c) {
                        int f[3] = {p[0], p[1], p[2]};
                        f[c] -= 1;
                        vf[c] += S(f[0], f[1], f[2]);
                        int g[3] = {p[0] - 1, p[1] - 1, p[2] - 1};
                        g[c] += 1;
                        ve[c] += S(g[0], g[1], g[2]);
                    }
                }
        const size_t at = size_t(leaf) * 512 + size_t(off);
        out[at] = vc > 0 ? vc * 0.125f : 0.f;
        for (int c = 0; c < 3; ++c) {
            out[size_t(1 + c) * n + at] = vf[c] > 0 ? vf[c] * 0.125f : 0.f;
            out[size_t(4 + c) * n + at] = ve[c] > 0 ? ve[c] * 0.125f : 0.f;
        }
    }
}

}  // namespace

hipError_t launch_visc_volumes(const VolumeStageArgs& a, int32_t nsub, int32_t nvol, float* sdf, float* svol,
                               float* out, hipStream_t s) {
    if (nsub > 0) {
        subcell_sdf_kernel<<<unsigned(nsub), kT, 0, s>>>(a, sdf);
        subcell_volume_kernel<<<unsigned(nsub), kT, 0, s>>>(a, sdf, svol);
    }
    if (nvol > 0) centre_volume_kernel<<<unsigned(nvol), kT, 0, s>>>(a, svol, out);
    return hipGetLastError();
}

}  // namespace uaamg
