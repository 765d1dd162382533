// Device kernels of the Poisson hierarchy build (hierarchy_device.h): level
// trimming with 1/diag, the coarse Galerkin sums, brick packing. One
// wavefront per leaf, one lane per z-row of 8 voxels (lane = x << 3 | y).
#include "hierarchy_device.h"
#include "stencil_device.h"

namespace uaamg {
namespace {

using namespace dev;

__device__ __forceinline__ uint32_t row_bits(const uint64_t* mask8, int lane) {
    return uint32_t(mask8[lane >> 3] >> ((lane & 7) * 8)) & 0xffu;
}

// largest value of the wavefront (values are never NaN: they start at 0 and
// are raised only by ordered comparisons)
__device__ __forceinline__ float wave_max(float m) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const float o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
    }
    return m;
}

__global__ __launch_bounds__(kBlockThreads) void level_trim_kernel(TrimArgs a) {
    const int leaf = wave_leaf();
    if (leaf >= a.n) return;
    const int lane = threadIdx.x & 63;
    const uint32_t rm = row_bits(a.mask + size_t(leaf) * 8, lane);
    const size_t src_leaf = a.gather ? size_t(a.gather[leaf]) : size_t(leaf);
    const size_t src_row = src_leaf * 512 + size_t(lane) * 8;
    const size_t row = size_t(leaf) * 512 + size_t(lane) * 8;
    float d[8];
    bool keep_diag = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float v[8];
        ld8(a.src[k] + src_row, v);
        // max over the level's DOF voxels of |v - default| (trimDefaultNodes)
        float m = 0.f;
#pragma unroll
        for (int z = 0; z < 8; ++z)
            if ((rm >> z) & 1u) {
                const float e = fabsf(v[z] - a.dflt[k]);
                if (e > m) m = e;
            }
        m = wave_max(m);
        const bool trimmed = m <= a.eps[k];
        if (trimmed)
#pragma unroll
            for (int z = 0; z < 8; ++z) v[z] = a.dflt[k];
        st8(a.dst[k] + row, v);
        if (lane == 0) a.keep[size_t(k) * size_t(a.n) + size_t(leaf)] = trimmed ? 0 : 1;
        if (k == 0) {
            keep_diag = !trimmed;
#pragma unroll
            for (int z = 0; z < 8; ++z) d[z] = v[z];
        }
    }
    float inv[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        if (!keep_diag || !((rm >> z) & 1u)) inv[z] = a.inv_dflt;
        else inv[z] = d[z] == 0.f ? 0.f : 1.0f / d[z];
    }
    st8(a.invdiag + row, inv);
}

__device__ __forceinline__ bool fine_on(const uint64_t* fmask, int32_t leaf, int lx, int ly, int lz) {
    if (leaf < 0) return false;
    const int off = (lx << 6) | (ly << 3) | lz;
    return (fmask[size_t(leaf) * 8 + size_t(off >> 6)] >> (off & 63)) & 1u;
}

__global__ __launch_bounds__(kBlockThreads) void galerkin_kernel(GalerkinArgs a) {
    const int leaf = wave_leaf();
    if (leaf >= a.n) return;
    const int lane = threadIdx.x & 63;
    const int cx = lane >> 3, cy = lane & 7;
    const uint32_t rm = row_bits(a.cmask + size_t(leaf) * 8, lane);
    const float t = a.term;
    float dg[8], ox[8], oy[8], oz[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        dg[z] = 6.0f * t;
        ox[z] = -t;
        oy[z] = -t;
        oz[z] = -t;
    }
    for (int cz = 0; cz < 8; ++cz) {
        if (!((rm >> cz) & 1u)) continue;
        // the 2^3 fine voxels lie in the child leaf of the voxel's octant; their
        // -x/-y/-z neighbours in that leaf or its face neighbour
        const int32_t fl = a.child[size_t(leaf) * 8 + size_t(((cx >= 4) << 2) | ((cy >= 4) << 1) | (cz >= 4))];
        const int32_t nx = fl >= 0 ? a.fnbr[size_t(fl) * 6 + 0] : -1;
        const int32_t ny = fl >= 0 ? a.fnbr[size_t(fl) * 6 + 2] : -1;
        const int32_t nz = fl >= 0 ? a.fnbr[size_t(fl) * 6 + 4] : -1;
        const int fx = 2 * (cx & 3), fy = 2 * (cy & 3), fz = 2 * (cz & 3);
        float diag = 0, x = 0, y = 0, zc = 0;
        for (int ii = 0; ii < 2; ++ii)
            for (int jj = 0; jj < 2; ++jj)
                for (int kk = 0; kk < 2; ++kk) {
                    const int lx = fx + ii, ly = fy + jj, lz = fz + kk;
                    if (!fine_on(a.fmask, fl, lx, ly, lz)) continue;
                    const size_t at = size_t(fl) * 512 + size_t((lx << 6) | (ly << 3) | lz);
                    diag += a.fine[0][at];
                    if (lx > 0 ? fine_on(a.fmask, fl, lx - 1, ly, lz) : fine_on(a.fmask, nx, 7, ly, lz)) {
                        if (ii == 0) x += a.fine[1][at];
                        else diag += 2 * a.fine[1][at];
                    }
                    if (ly > 0 ? fine_on(a.fmask, fl, lx, ly - 1, lz) : fine_on(a.fmask, ny, lx, 7, lz)) {
                        if (jj == 0) y += a.fine[2][at];
                        else diag += 2 * a.fine[2][at];
                    }
                    if (lz > 0 ? fine_on(a.fmask, fl, lx, ly, lz - 1) : fine_on(a.fmask, nz, lx, ly, 7)) {
                        if (kk == 0) zc += a.fine[3][at];
                        else diag += 2 * a.fine[3][at];
                    }
                }
        const float factor = 0.5f * (1.0f / 8.0f);
        dg[cz] = diag * factor;
        ox[cz] = x * factor;
        oy[cz] = y * factor;
        oz[cz] = zc * factor;
    }
    const size_t row = size_t(leaf) * 512 + size_t(lane) * 8;
    st8(a.coarse[0] + row, dg);
    st8(a.coarse[1] + row, ox);
    st8(a.coarse[2] + row, oy);
    st8(a.coarse[3] + row, oz);
}

__global__ __launch_bounds__(kBlockThreads) void pack_bricks_kernel(int32_t n, const int32_t* slot, const float* src,
                                                                    float* dst) {
    const int leaf = wave_leaf();
    if (leaf >= n) return;
    const int32_t s = slot[leaf];
    if (s < 0) return;
    const int lane = threadIdx.x & 63;
    float v[8];
    ld8(src + size_t(leaf) * 512 + size_t(lane) * 8, v);
    st8(dst + size_t(s) * 512 + size_t(lane) * 8, v);
}

}  // namespace

hipError_t launch_level_trim(const TrimArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    level_trim_kernel<<<grid_for(a.n), kBlockThreads, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_galerkin(const GalerkinArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    galerkin_kernel<<<grid_for(a.n), kBlockThreads, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_pack_bricks(int32_t n, const int32_t* slot, const float* src, float* dst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    pack_bricks_kernel<<<grid_for(n), kBlockThreads, 0, s>>>(n, slot, src, dst);
    return hipGetLastError();
}

}  // namespace uaamg
