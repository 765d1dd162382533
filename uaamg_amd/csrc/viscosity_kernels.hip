// Variational viscosity kernels: the 15-point staggered operator in every
// applier mode, the transfers between levels, the coarsest Eigen-CG
// restatement. Arithmetic per value follows light_weight_applier::
// component_operator<c> (simd_viscosity3d.cpp:2841-3582): products rounded one
// by one, the same-channel sum ((x+ + x-) + (y+ + y-)) + (z+ + z-), then the two
// cross-channel groups, fused multiply-adds exactly where the AVX code uses
// _mm256_fmadd_ps (build with -ffp-contract=off).
#include <cfloat>

#include "stencil_device.h"
#include "viscosity_host.h"
#include "viscosity_kernels.h"
#include "wave_reduce.h"
#include "xcd_map.h"

namespace uaamg {
namespace {

constexpr int kVWaves = 4;
constexpr int kVThreads = 64 * kVWaves;

constexpr int nb_slot(int ax, int ay, int az) { return (ax + 1) * 9 + (ay + 1) * 3 + (az + 1); }

__device__ __forceinline__ int row_off(int X, int Y) { return ((X & 7) * 8 + (Y & 7)) * 8; }

__device__ __forceinline__ void ld8v(const float* p, float v[8]) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    const float4 b = *reinterpret_cast<const float4*>(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void st8v(float* p, const float v[8]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
}
// BF16-stored level vectors of the mixed form: a z-row is 16 bytes, widened
// exactly on load, rounded to nearest even on store; arithmetic stays FP32
__device__ __forceinline__ void ld8v(const bf16* p, float v[8]) { dev::ld8(p, v); }
__device__ __forceinline__ void st8v(bf16* p, const float v[8]) { dev::st8(p, v); }
__device__ __forceinline__ float ld1v(const float* p) { return *p; }
__device__ __forceinline__ float ld1v(const bf16* p) { return float(*p); }
__device__ __forceinline__ void st1v(float* p, float v) { *p = v; }
__device__ __forceinline__ void st1v(bf16* p, float v) { *p = bf16(v); }

// A wave record (layout at kVRecLeaf) in scalar registers.
struct VRec {
    int leaf;
    int nb[27];
    int own[14];
    int nbs[7];
};

__device__ __forceinline__ VRec load_vrec(const ViscLevelView& L, int e) {
    const int4* p = reinterpret_cast<const int4*>(L.wrec + size_t(e) * kVRec);
    int v[52];
#pragma unroll
    for (int i = 0; i < 13; ++i) {
        const int4 q = p[i];
        v[4 * i] = q.x; v[4 * i + 1] = q.y; v[4 * i + 2] = q.z; v[4 * i + 3] = q.w;
    }
    VRec r;
    r.leaf = v[kVRecLeaf];
#pragma unroll
    for (int k = 0; k < 27; ++k) r.nb[k] = v[kVRecNb + k];
#pragma unroll
    for (int k = 0; k < 14; ++k) r.own[k] = v[kVRecOwn + k];
#pragma unroll
    for (int k = 0; k < 7; ++k) r.nbs[k] = v[kVRecNbr + k];
    // One use of every field here, in the kernel's first basic block: the
    // record's scalar loads are then issued together and waited for once
    // (otherwise each is issued in the branch that first reads it). The
    // empty asm emits no instruction.
    int all = 0;
#pragma unroll
    for (int k = 0; k < 49; ++k) all ^= v[k];
    asm volatile("" ::"s"(all));
    return r;
}

// Leaf holding leaf-local row (lx + OX, ly + OY), shifted AZ leaves in z:
// one of four record entries chosen per lane.
template <int OX, int OY, int AZ>
__device__ __forceinline__ int vleaf(const VRec& R, int lx, int ly) {
    constexpr int sx = (OX > 0) - (OX < 0), sy = (OY > 0) - (OY < 0);
    const int X = lx + OX, Y = ly + OY;
    const bool ox = X < 0 || X > 7, oy = Y < 0 || Y > 7;
    const int l00 = R.nb[nb_slot(0, 0, AZ)], l10 = R.nb[nb_slot(sx, 0, AZ)];
    const int l01 = R.nb[nb_slot(0, sy, AZ)], l11 = R.nb[nb_slot(sx, sy, AZ)];
    return ox ? (oy ? l11 : l10) : (oy ? l01 : l00);
}

// channel ch of vector v at (lx + OX, ly + OY, z + OZ), z = 0..7 (a missing
// leaf, or one without vector bricks, reads 0)
template <int OX, int OY, int OZ, class V>
__device__ __forceinline__ void vfetch(const ViscLevelView& L, const V* v, int ch, const VRec& R, int lx, int ly,
                                       float out[8]) {
    const int X = lx + OX, Y = ly + OY;
    const int lf = vleaf<OX, OY, 0>(R, lx, ly);
    float row[8];
    if (lf >= 0 && lf < L.nleaf_dof) ld8v(v + (size_t(lf) * 3 + ch) * 512 + row_off(X, Y), row);
    else {
#pragma unroll
        for (int k = 0; k < 8; ++k) row[k] = 0.f;
    }
    if (OZ == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) out[k] = row[k];
    } else if (OZ > 0) {
#pragma unroll
        for (int k = 0; k < 7; ++k) out[k] = row[k + 1];
        const int lz = vleaf<OX, OY, 1>(R, lx, ly);
        out[7] = (lz >= 0 && lz < L.nleaf_dof) ? ld1v(v + (size_t(lz) * 3 + ch) * 512 + row_off(X, Y)) : 0.f;
    } else {
#pragma unroll
        for (int k = 7; k > 0; --k) out[k] = row[k - 1];
        const int lz = vleaf<OX, OY, -1>(R, lx, ly);
        out[0] = (lz >= 0 && lz < L.nleaf_dof) ? ld1v(v + (size_t(lz) * 3 + ch) * 512 + row_off(X, Y) + 7) : 0.f;
    }
}

// coefficient grid GID at (lx + OX, ly + OY, z + OZ), at most one offset
// nonzero: rows inside the leaf read slot s_own, rows (or the shifted z
// value) beyond the face read s_nb; slot -1 reads the level default
template <int GID, int OX, int OY, int OZ>
__device__ __forceinline__ void cfetch(const ViscLevelView& L, int s_own, int s_nb, int lx, int ly, float out[8]) {
    static_assert((OX != 0) + (OY != 0) + (OZ != 0) <= 1, "single-axis coefficient offsets");
    const int X = lx + OX, Y = ly + OY;
    const bool beyond = X < 0 || X > 7 || Y < 0 || Y > 7;
    const int s = beyond ? s_nb : s_own;
    const float m = L.dflt[GID];
    float row[8];
    if (s >= 0) ld8v(L.pool + size_t(s) * 512 + row_off(X, Y), row);
    else {
#pragma unroll
        for (int k = 0; k < 8; ++k) row[k] = m;
    }
    if (OZ == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k) out[k] = row[k];
    } else if (OZ > 0) {
#pragma unroll
        for (int k = 0; k < 7; ++k) out[k] = row[k + 1];
        out[7] = s_nb >= 0 ? L.pool[size_t(s_nb) * 512 + row_off(X, Y)] : m;
    } else {
#pragma unroll
        for (int k = 7; k > 0; --k) out[k] = row[k - 1];
        out[0] = s_nb >= 0 ? L.pool[size_t(s_nb) * 512 + row_off(X, Y) + 7] : m;
    }
}

template <int A>
struct Unit {
    static constexpr int x = A == 0, y = A == 1, z = A == 2;
};

// One row (lane) of the applier on channel C for wave record e (a DOF leaf
// of channel C): Y = op(X, B). Every load is issued unconditionally (all in
// bounds); rows without a DOF skip only the store.
template <int MODE, int C, class V>
__device__ __forceinline__ void visc_row(const ViscLevelView& L, int e, int lane, const V* Xv, const V* Bv, V* Yv) {
    const int lx = lane >> 3, ly = lane & 7;
    const VRec R = load_vrec(L, e);
    const int leaf = R.leaf;
    const uint32_t rm = uint32_t(L.mask[(size_t(leaf) * 3 + C) * 8 + lx] >> (ly * 8)) & 0xffu;
    constexpr int d0 = C, d1 = (C + 1) % 3, d2 = (C + 2) % 3;
    using E0 = Unit<d0>;
    using E1 = Unit<d1>;
    using E2 = Unit<d2>;
    constexpr int kN0 = kVNeg + 3 * C, kX2 = kVCross + 4 * d2, kX1 = kVCross + 4 * d1;

    float s[8];
    {
        float a[8], wa[8], b[8], wb[8];
        float f12[8], f34[8];
        vfetch<1, 0, 0>(L, Xv, C, R, lx, ly, a);
        cfetch<kN0 + 0, 1, 0, 0>(L, R.own[0], R.nbs[0], lx, ly, wa);
        vfetch<-1, 0, 0>(L, Xv, C, R, lx, ly, b);
        cfetch<kN0 + 0, 0, 0, 0>(L, R.own[0], -1, lx, ly, wb);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            f12[k] = p + q;
        }
        vfetch<0, 1, 0>(L, Xv, C, R, lx, ly, a);
        cfetch<kN0 + 1, 0, 1, 0>(L, R.own[1], R.nbs[1], lx, ly, wa);
        vfetch<0, -1, 0>(L, Xv, C, R, lx, ly, b);
        cfetch<kN0 + 1, 0, 0, 0>(L, R.own[1], -1, lx, ly, wb);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            f34[k] = p + q;
        }
        vfetch<0, 0, 1>(L, Xv, C, R, lx, ly, a);
        cfetch<kN0 + 2, 0, 0, 1>(L, R.own[2], R.nbs[2], lx, ly, wa);
        vfetch<0, 0, -1>(L, Xv, C, R, lx, ly, b);
        cfetch<kN0 + 2, 0, 0, 0>(L, R.own[2], -1, lx, ly, wb);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            const float f56 = p + q;
            const float t = f12[k] + f34[k];
            s[k] = t + f56;
        }
    }
    {
        // normal d2: channel d1 at v, v + e1, v - e0, v - e0 + e1
        float u[8], c[8], t01[8];
        vfetch<0, 0, 0>(L, Xv, d1, R, lx, ly, u);
        cfetch<kX2 + 0, 0, 0, 0>(L, R.own[3], -1, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) t01[k] = u[k] * c[k];
        vfetch<E1::x, E1::y, E1::z>(L, Xv, d1, R, lx, ly, u);
        cfetch<kX2 + 1, 0, 0, 0>(L, R.own[4], -1, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = c[k] * u[k];
            t01[k] = t01[k] + p;
        }
        float t23[8];
        vfetch<-E0::x, -E0::y, -E0::z>(L, Xv, d1, R, lx, ly, u);
        cfetch<kX2 + 2, -E0::x, -E0::y, -E0::z>(L, R.own[5], R.nbs[3], lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) t23[k] = c[k] * u[k];
        vfetch<E1::x - E0::x, E1::y - E0::y, E1::z - E0::z>(L, Xv, d1, R, lx, ly, u);
        cfetch<kX2 + 3, -E0::x, -E0::y, -E0::z>(L, R.own[6], R.nbs[4], lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = c[k] * u[k];
            t23[k] = t23[k] + p;
            const float T2 = t01[k] + t23[k];
            s[k] = s[k] + T2;
        }
    }
    {
        // normal d1: channel d2 at v, v - e0, v + e2, v - e0 + e2
        float u[8], c[8], t01[8];
        vfetch<0, 0, 0>(L, Xv, d2, R, lx, ly, u);
        cfetch<kX1 + 0, 0, 0, 0>(L, R.own[7], -1, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) t01[k] = c[k] * u[k];
        vfetch<-E0::x, -E0::y, -E0::z>(L, Xv, d2, R, lx, ly, u);
        cfetch<kX1 + 1, -E0::x, -E0::y, -E0::z>(L, R.own[8], R.nbs[5], lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = c[k] * u[k];
            t01[k] = t01[k] + p;
        }
        float t23[8];
        vfetch<E2::x, E2::y, E2::z>(L, Xv, d2, R, lx, ly, u);
        cfetch<kX1 + 2, 0, 0, 0>(L, R.own[9], -1, lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) t23[k] = c[k] * u[k];
        vfetch<E2::x - E0::x, E2::y - E0::y, E2::z - E0::z>(L, Xv, d2, R, lx, ly, u);
        cfetch<kX1 + 3, -E0::x, -E0::y, -E0::z>(L, R.own[10], R.nbs[6], lx, ly, c);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float p = c[k] * u[k];
            t23[k] = t23[k] + p;
            const float T1 = t01[k] + t23[k];
            s[k] = s[k] + T1;
        }
    }

    const size_t row = (size_t(leaf) * 3 + C) * 512 + size_t(lx * 8 + ly) * 8;
    float lhs[8], y[8];
    ld8v(Xv + row, lhs);
    if (MODE == kVRedGS || MODE == kVBlackGS) {
        float b[8], inv[8];
        ld8v(Bv + row, b);
        cfetch<kVInvDiag + C, 0, 0, 0>(L, R.own[12], -1, lx, ly, inv);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float t = b[k] - s[k];
            t = t * inv[k];
            t = t * 1.0f;
            const float r = fmaf(lhs[k], 0.0f, t);
            const bool red = ((lx + ly + k) & 1) == 0;
            y[k] = (red == (MODE == kVRedGS)) ? r : lhs[k];
        }
    } else {
        float dg[8];
        cfetch<kVDiag + C, 0, 0, 0>(L, R.own[11], -1, lx, ly, dg);
        if (MODE == kVNormal) {
#pragma unroll
            for (int k = 0; k < 8; ++k) y[k] = fmaf(lhs[k], dg[k], s[k]);
        } else {
            float b[8];
            ld8v(Bv + row, b);
            if (MODE == kVResidual) {
#pragma unroll
                for (int k = 0; k < 8; ++k) y[k] = b[k] - fmaf(lhs[k], dg[k], s[k]);
            } else {
                float sc[8];
                if (MODE == kVJacobi) cfetch<kVInvDiag + C, 0, 0, 0>(L, R.own[12], -1, lx, ly, sc);
                else cfetch<kVSpai0 + C, 0, 0, 0>(L, R.own[13], -1, lx, ly, sc);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const float a = fmaf(lhs[k], dg[k], s[k]);
                    float t = b[k] - a;
                    t = t * sc[k];
                    y[k] = fmaf(1.0f, t, lhs[k]);
                }
            }
        }
    }
    if (!rm) return;   // rows without a DOF are not written
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (!((rm >> k) & 1u)) y[k] = 0.f;
    st8v(Yv + row, y);
}

// ---------------------------------------------------------- part rows
// The same row arithmetic on part of a z-row per lane: 8 / NZ wavefronts
// share a (leaf, channel) record, wavefront q taking z = NZ*q .. NZ*q + NZ-1
// of every row (NZ = 4: half rows, 2: quarter rows). A lane then needs
// NZ / 8 of a full row's operands, so a wavefront's loads fit in flight at
// once instead of in batches; per voxel the terms, their order and the
// stored values are those of visc_row.
template <int NZ>
__device__ __forceinline__ void ldnv(const float* p, float v[NZ]) {
    if (NZ == 4) {
        const float4 a = *reinterpret_cast<const float4*>(p);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else {
        const float2 a = *reinterpret_cast<const float2*>(p);
        v[0] = a.x; v[1] = a.y;
    }
}
template <int NZ>
__device__ __forceinline__ void stnv(float* p, const float v[NZ]) {
    if (NZ == 4) *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    else *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
}
template <int NZ>
__device__ __forceinline__ void ldnv(const bf16* p, float v[NZ]) {
    if (NZ == 4) {
        const float4 a = dev::ld4(p);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else {
        const uint32_t q = *reinterpret_cast<const uint32_t*>(p);
        v[0] = __uint_as_float(q << 16);
        v[1] = __uint_as_float(q & 0xffff0000u);
    }
}
template <int NZ>
__device__ __forceinline__ void stnv(bf16* p, const float v[NZ]) {
    if (NZ == 4) *reinterpret_cast<uint2*>(p) = make_uint2(dev::pack_bf16(v[0], v[1]), dev::pack_bf16(v[2], v[3]));
    else *reinterpret_cast<uint32_t*>(p) = dev::pack_bf16(v[0], v[1]);
}

// vfetch on part of a row: z = z0 .. z0 + NZ - 1 (z0 the same for the whole
// wavefront); the value a z shift brings in comes from the own row or, past
// the brick, from the +-z neighbour leaf
template <int OX, int OY, int OZ, int NZ, class V>
__device__ __forceinline__ void vfetchn(const ViscLevelView& L, const V* v, int ch, const VRec& R, int lx, int ly,
                                        int z0, float out[NZ]) {
    const int X = lx + OX, Y = ly + OY;
    const int lf = vleaf<OX, OY, 0>(R, lx, ly);
    const bool have = lf >= 0 && lf < L.nleaf_dof;
    const V* base = v + (size_t(have ? lf : 0) * 3 + ch) * 512 + row_off(X, Y);
    float row[NZ];
    if (have) ldnv<NZ>(base + z0, row);
    else {
#pragma unroll
        for (int k = 0; k < NZ; ++k) row[k] = 0.f;
    }
    if (OZ == 0) {
#pragma unroll
        for (int k = 0; k < NZ; ++k) out[k] = row[k];
    } else if (OZ > 0) {
#pragma unroll
        for (int k = 0; k < NZ - 1; ++k) out[k] = row[k + 1];
        if (z0 + NZ <= 7) {
            out[NZ - 1] = have ? ld1v(base + z0 + NZ) : 0.f;
        } else {
            const int lz = vleaf<OX, OY, 1>(R, lx, ly);
            out[NZ - 1] =
                (lz >= 0 && lz < L.nleaf_dof) ? ld1v(v + (size_t(lz) * 3 + ch) * 512 + row_off(X, Y)) : 0.f;
        }
    } else {
#pragma unroll
        for (int k = NZ - 1; k > 0; --k) out[k] = row[k - 1];
        if (z0 > 0) {
            out[0] = have ? ld1v(base + z0 - 1) : 0.f;
        } else {
            const int lz = vleaf<OX, OY, -1>(R, lx, ly);
            out[0] = (lz >= 0 && lz < L.nleaf_dof) ? ld1v(v + (size_t(lz) * 3 + ch) * 512 + row_off(X, Y) + 7) : 0.f;
        }
    }
}

// cfetch on part of a row (same rules as cfetch)
template <int GID, int OX, int OY, int OZ, int NZ>
__device__ __forceinline__ void cfetchn(const ViscLevelView& L, int s_own, int s_nb, int lx, int ly, int z0,
                                        float out[NZ]) {
    static_assert((OX != 0) + (OY != 0) + (OZ != 0) <= 1, "single-axis coefficient offsets");
    const int X = lx + OX, Y = ly + OY;
    const bool beyond = X < 0 || X > 7 || Y < 0 || Y > 7;
    const int s = beyond ? s_nb : s_own;
    const float m = L.dflt[GID];
    float row[NZ];
    if (s >= 0) ldnv<NZ>(L.pool + size_t(s) * 512 + row_off(X, Y) + z0, row);
    else {
#pragma unroll
        for (int k = 0; k < NZ; ++k) row[k] = m;
    }
    if (OZ == 0) {
#pragma unroll
        for (int k = 0; k < NZ; ++k) out[k] = row[k];
    } else if (OZ > 0) {
#pragma unroll
        for (int k = 0; k < NZ - 1; ++k) out[k] = row[k + 1];
        if (z0 + NZ <= 7) out[NZ - 1] = s >= 0 ? L.pool[size_t(s) * 512 + row_off(X, Y) + z0 + NZ] : m;
        else out[NZ - 1] = s_nb >= 0 ? L.pool[size_t(s_nb) * 512 + row_off(X, Y)] : m;
    } else {
#pragma unroll
        for (int k = NZ - 1; k > 0; --k) out[k] = row[k - 1];
        if (z0 > 0) out[0] = s >= 0 ? L.pool[size_t(s) * 512 + row_off(X, Y) + z0 - 1] : m;
        else out[0] = s_nb >= 0 ? L.pool[size_t(s_nb) * 512 + row_off(X, Y) + 7] : m;
    }
}

template <int MODE, int C, int NZ, class V>
__device__ __forceinline__ void visc_rown(const ViscLevelView& L, int e, int lane, int z0, const V* Xv, const V* Bv,
                                          V* Yv) {
    const int lx = lane >> 3, ly = lane & 7;
    const VRec R = load_vrec(L, e);
    const int leaf = R.leaf;
    const uint32_t rm = uint32_t(L.mask[(size_t(leaf) * 3 + C) * 8 + lx] >> (ly * 8)) & 0xffu;
    constexpr int d0 = C, d1 = (C + 1) % 3, d2 = (C + 2) % 3;
    using E0 = Unit<d0>;
    using E1 = Unit<d1>;
    using E2 = Unit<d2>;
    constexpr int kN0 = kVNeg + 3 * C, kX2 = kVCross + 4 * d2, kX1 = kVCross + 4 * d1;

    float s[NZ];
    {
        float a[NZ], wa[NZ], b[NZ], wb[NZ];
        float f12[NZ], f34[NZ];
        vfetchn<1, 0, 0, NZ>(L, Xv, C, R, lx, ly, z0, a);
        cfetchn<kN0 + 0, 1, 0, 0, NZ>(L, R.own[0], R.nbs[0], lx, ly, z0, wa);
        vfetchn<-1, 0, 0, NZ>(L, Xv, C, R, lx, ly, z0, b);
        cfetchn<kN0 + 0, 0, 0, 0, NZ>(L, R.own[0], -1, lx, ly, z0, wb);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            f12[k] = p + q;
        }
        vfetchn<0, 1, 0, NZ>(L, Xv, C, R, lx, ly, z0, a);
        cfetchn<kN0 + 1, 0, 1, 0, NZ>(L, R.own[1], R.nbs[1], lx, ly, z0, wa);
        vfetchn<0, -1, 0, NZ>(L, Xv, C, R, lx, ly, z0, b);
        cfetchn<kN0 + 1, 0, 0, 0, NZ>(L, R.own[1], -1, lx, ly, z0, wb);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            f34[k] = p + q;
        }
        vfetchn<0, 0, 1, NZ>(L, Xv, C, R, lx, ly, z0, a);
        cfetchn<kN0 + 2, 0, 0, 1, NZ>(L, R.own[2], R.nbs[2], lx, ly, z0, wa);
        vfetchn<0, 0, -1, NZ>(L, Xv, C, R, lx, ly, z0, b);
        cfetchn<kN0 + 2, 0, 0, 0, NZ>(L, R.own[2], -1, lx, ly, z0, wb);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const float p = a[k] * wa[k], q = b[k] * wb[k];
            const float f56 = p + q;
            const float t = f12[k] + f34[k];
            s[k] = t + f56;
        }
    }
    {
        // normal d2: channel d1 at v, v + e1, v - e0, v - e0 + e1
        float u[NZ], c[NZ], t01[NZ];
        vfetchn<0, 0, 0, NZ>(L, Xv, d1, R, lx, ly, z0, u);
        cfetchn<kX2 + 0, 0, 0, 0, NZ>(L, R.own[3], -1, lx, ly, z0, c);
#pragma unroll
        for (int k = 0; k < NZ; ++k) t01[k] = u[k] * c[k];
        vfetchn<E1::x, E1::y, E1::z, NZ>(L, Xv, d1, R, lx, ly, z0, u);
        cfetchn<kX2 + 1, 0, 0, 0, NZ>(L, R.own[4], -1, lx, ly, z0, c);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const float p = c[k] * u[k];
            t01[k] = t01[k] + p;
        }
        float t23[NZ];
        vfetchn<-E0::x, -E0::y, -E0::z, NZ>(L, Xv, d1, R, lx, ly, z0, u);
        cfetchn<kX2 + 2, -E0::x, -E0::y, -E0::z, NZ>(L, R.own[5], R.nbs[3], lx, ly, z0, c);
#pragma unroll
        for (int k = 0; k < NZ; ++k) t23[k] = c[k] * u[k];
        vfetchn<E1::x - E0::x, E1::y - E0::y, E1::z - E0::z, NZ>(L, Xv, d1, R, lx, ly, z0, u);
        cfetchn<kX2 + 3, -E0::x, -E0::y, -E0::z, NZ>(L, R.own[6], R.nbs[4], lx, ly, z0, c);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const float p = c[k] * u[k];
            t23[k] = t23[k] + p;
            const float T2 = t01[k] + t23[k];
            s[k] = s[k] + T2;
        }
    }
    {
        // normal d1: channel d2 at v, v - e0, v + e2, v - e0 + e2
        float u[NZ], c[NZ], t01[NZ];
        vfetchn<0, 0, 0, NZ>(L, Xv, d2, R, lx, ly, z0, u);
        cfetchn<kX1 + 0, 0, 0, 0, NZ>(L, R.own[7], -1, lx, ly, z0, c);
#pragma unroll
        for (int k = 0; k < NZ; ++k) t01[k] = c[k] * u[k];
        vfetchn<-E0::x, -E0::y, -E0::z, NZ>(L, Xv, d2, R, lx, ly, z0, u);
        cfetchn<kX1 + 1, -E0::x, -E0::y, -E0::z, NZ>(L, R.own[8], R.nbs[5], lx, ly, z0, c);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const float p = c[k] * u[k];
            t01[k] = t01[k] + p;
        }
        float t23[NZ];
        vfetchn<E2::x, E2::y, E2::z, NZ>(L, Xv, d2, R, lx, ly, z0, u);
        cfetchn<kX1 + 2, 0, 0, 0, NZ>(L, R.own[9], -1, lx, ly, z0, c);
#pragma unroll
        for (int k = 0; k < NZ; ++k) t23[k] = c[k] * u[k];
        vfetchn<E2::x - E0::x, E2::y - E0::y, E2::z - E0::z, NZ>(L, Xv, d2, R, lx, ly, z0, u);
        cfetchn<kX1 + 3, -E0::x, -E0::y, -E0::z, NZ>(L, R.own[10], R.nbs[6], lx, ly, z0, c);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const float p = c[k] * u[k];
            t23[k] = t23[k] + p;
            const float T1 = t01[k] + t23[k];
            s[k] = s[k] + T1;
        }
    }

    const size_t row = (size_t(leaf) * 3 + C) * 512 + size_t(lx * 8 + ly) * 8 + size_t(z0);
    float lhs[NZ], y[NZ];
    ldnv<NZ>(Xv + row, lhs);
    if (MODE == kVRedGS || MODE == kVBlackGS) {
        float b[NZ], inv[NZ];
        ldnv<NZ>(Bv + row, b);
        cfetchn<kVInvDiag + C, 0, 0, 0, NZ>(L, R.own[12], -1, lx, ly, z0, inv);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            float t = b[k] - s[k];
            t = t * inv[k];
            t = t * 1.0f;
            const float r = fmaf(lhs[k], 0.0f, t);
            const bool red = ((lx + ly + z0 + k) & 1) == 0;
            y[k] = (red == (MODE == kVRedGS)) ? r : lhs[k];
        }
    } else {
        float dg[NZ];
        cfetchn<kVDiag + C, 0, 0, 0, NZ>(L, R.own[11], -1, lx, ly, z0, dg);
        if (MODE == kVNormal) {
#pragma unroll
            for (int k = 0; k < NZ; ++k) y[k] = fmaf(lhs[k], dg[k], s[k]);
        } else {
            float b[NZ];
            ldnv<NZ>(Bv + row, b);
            if (MODE == kVResidual) {
#pragma unroll
                for (int k = 0; k < NZ; ++k) y[k] = b[k] - fmaf(lhs[k], dg[k], s[k]);
            } else {
                float sc[NZ];
                if (MODE == kVJacobi) cfetchn<kVInvDiag + C, 0, 0, 0, NZ>(L, R.own[12], -1, lx, ly, z0, sc);
                else cfetchn<kVSpai0 + C, 0, 0, 0, NZ>(L, R.own[13], -1, lx, ly, z0, sc);
#pragma unroll
                for (int k = 0; k < NZ; ++k) {
                    const float a = fmaf(lhs[k], dg[k], s[k]);
                    float t = b[k] - a;
                    t = t * sc[k];
                    y[k] = fmaf(1.0f, t, lhs[k]);
                }
            }
        }
    }
    const uint32_t rmh = (rm >> z0) & ((1u << NZ) - 1u);
    if (!rmh) return;   // row parts without a DOF are not written
#pragma unroll
    for (int k = 0; k < NZ; ++k)
        if (!((rmh >> k) & 1u)) y[k] = 0.f;
    stnv<NZ>(Yv + row, y);
}


template <int MODE, int NZ, class V>
__global__ __launch_bounds__(kVThreads) void visc_apply_all_part_kernel(ViscLevelView L, const V* Xv, const V* Bv,
                                                                        V* Yv) {
    constexpr int kParts = 8 / NZ;
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    const int e = w / kParts, z0 = (w % kParts) * NZ;
    const int lane = threadIdx.x & 63;
    if (e < L.chan_off[1]) visc_rown<MODE, 0, NZ>(L, e, lane, z0, Xv, Bv, Yv);
    else if (e < L.chan_off[2]) visc_rown<MODE, 1, NZ>(L, e, lane, z0, Xv, Bv, Yv);
    else if (e < L.chan_off[3]) visc_rown<MODE, 2, NZ>(L, e, lane, z0, Xv, Bv, Yv);
}

template <int MODE, int C, int NZ, class V>
__global__ __launch_bounds__(kVThreads) void visc_apply_part_kernel(ViscLevelView L, const V* Xv, const V* Bv, V* Yv) {
    constexpr int kParts = 8 / NZ;
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    const int e = w / kParts;
    if (e >= L.chan_off[C + 1] - L.chan_off[C]) return;
    visc_rown<MODE, C, NZ>(L, L.chan_off[C] + e, threadIdx.x & 63, (w % kParts) * NZ, Xv, Bv, Yv);
}

// All three channels in one launch (modes that write y != x: the channels
// are independent); the wave's record index selects the channel.
template <int MODE, class V>
__global__ __launch_bounds__(kVThreads) void visc_apply_all_kernel(ViscLevelView L, const V* Xv, const V* Bv, V* Yv) {
    const int e = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (e < L.chan_off[1]) visc_row<MODE, 0>(L, e, lane, Xv, Bv, Yv);
    else if (e < L.chan_off[2]) visc_row<MODE, 1>(L, e, lane, Xv, Bv, Yv);
    else if (e < L.chan_off[3]) visc_row<MODE, 2>(L, e, lane, Xv, Bv, Yv);
}

template <int MODE, int C, class V>
__global__ __launch_bounds__(kVThreads) void visc_apply_kernel(ViscLevelView L, const V* Xv, const V* Bv, V* Yv) {
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    if (w >= L.chan_off[C + 1] - L.chan_off[C]) return;
    visc_row<MODE, C>(L, L.chan_off[C] + w, threadIdx.x & 63, Xv, Bv, Yv);
}

// The same two kernels for levels with many leaves (L.wide): two waves per
// SIMD (the register budget halves; the compiler keeps fewer of a row's loads
// in flight per wave, the second wave hides their latency). Same arithmetic,
// same results.
template <int MODE, class V>
__global__ __launch_bounds__(kVThreads) __attribute__((amdgpu_waves_per_eu(2))) void visc_apply_all_wide_kernel(
    ViscLevelView L, const V* Xv, const V* Bv, V* Yv) {
    const int e = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (e < L.chan_off[1]) visc_row<MODE, 0>(L, e, lane, Xv, Bv, Yv);
    else if (e < L.chan_off[2]) visc_row<MODE, 1>(L, e, lane, Xv, Bv, Yv);
    else if (e < L.chan_off[3]) visc_row<MODE, 2>(L, e, lane, Xv, Bv, Yv);
}

template <int MODE, int C, class V>
__global__ __launch_bounds__(kVThreads) __attribute__((amdgpu_waves_per_eu(2))) void visc_apply_wide_kernel(
    ViscLevelView L, const V* Xv, const V* Bv, V* Yv) {
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    if (w >= L.chan_off[C + 1] - L.chan_off[C]) return;
    visc_row<MODE, C>(L, L.chan_off[C] + w, threadIdx.x & 63, Xv, Bv, Yv);
}

// restriction of one (coarse leaf, channel) row: 1/8 of the on fine children.
// The row's two children (z < 4 and z >= 4) and their rows are loaded
// unconditionally (an absent child reads leaf 0 and is dropped), so the
// loads depend on one round trip; sums run in the reference's order.
template <class VF, class VC>
__device__ __forceinline__ void visc_restrict_row(const ViscLevelView& Cl, const ViscLevelView& F, int leaf, int c,
                                                  int lane, const VF* fine, VC* coarse) {
    const int lx = lane >> 3, ly = lane & 7;
    const uint32_t rm = uint32_t(Cl.mask[(size_t(leaf) * 3 + c) * 8 + lx] >> (ly * 8)) & 0xffu;
    const int fx = 2 * (lx & 3), fy = 2 * (ly & 3);
    int f[2];
    float v[2][2][2][8];      // [child][a][b][fine z]
    uint64_t mw[2][2];        // [child][a] fine mask word of fine x = fx + a
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int oct = ((lx >> 2) << 2) | ((ly >> 2) << 1) | h;
        f[h] = Cl.child[size_t(leaf) * 8 + oct];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const size_t fl = size_t(f[h] >= 0 ? f[h] : 0);
        const uint64_t* fm = F.mask + (fl * 3 + c) * 8;
        const VF* fv = fine + (fl * 3 + c) * 512;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            mw[h][a] = fm[fx + a];
#pragma unroll
            for (int b = 0; b < 2; ++b) ld8v(fv + (fx + a) * 64 + (fy + b) * 8, v[h][a][b]);
        }
    }
    float out[8];
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const int h = z >> 2;
        const int fz = 2 * (z & 3);
        float sum = 0.f;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    const int bit = (fy + b) * 8 + fz + d;
                    if ((mw[h][a] >> bit) & 1u) sum += v[h][a][b][fz + d];
                }
        const float o = sum * 0.125f * 0.5f * 2.0f;
        out[z] = (((rm >> z) & 1u) && f[h] >= 0) ? o : 0.f;
    }
    if (!rm) return;   // rows without a DOF are not written
    st8v(coarse + (size_t(leaf) * 3 + c) * 512 + size_t(lx * 8 + ly) * 8, out);
}

// prolongation<true> of one (fine leaf, channel) row
template <class VF, class VC>
__device__ __forceinline__ void visc_prolong_row(const ViscLevelView& F, int leaf, int c, int lane, const VC* coarse,
                                                 VF* fine) {
    const int lx = lane >> 3, ly = lane & 7;
    const uint32_t rm = uint32_t(F.mask[(size_t(leaf) * 3 + c) * 8 + lx] >> (ly * 8)) & 0xffu;
    const int p = F.parent[leaf];
    if (p < 0 || !rm) return;
    const int cl = p >> 3, oct = p & 7;
    const int cx = ((oct >> 2) & 1) * 4 + (lx >> 1), cy = ((oct >> 1) & 1) * 4 + (ly >> 1), cz0 = (oct & 1) * 4;
    const VC* cv = coarse + (size_t(cl) * 3 + c) * 512 + cx * 64 + cy * 8 + cz0;
    VF* row = fine + (size_t(leaf) * 3 + c) * 512 + size_t(lx * 8 + ly) * 8;
    float v[8];
    ld8v(row, v);
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if ((rm >> z) & 1u) v[z] += ld1v(cv + (z >> 1));
    st8v(row, v);
}

// ---------------------------------------------------- small-level visits
// On the coarse levels a W-cycle visit is ~29 dependent launches (12 colour
// passes, residual, restriction, prolongation, 12 colour passes). When every
// channel of a level has at most a few dozen leaves, one workgroup runs the
// whole half-visit with a workgroup barrier between passes: the same row
// routine, the same pass order, the same values.
constexpr int kSmallThreads = 256;
constexpr int kSmallWaves = kSmallThreads / 64;

template <int MODE, int C, class V>
__device__ __forceinline__ void small_pass(const ViscLevelView& L, const V* X, const V* B, V* Y) {
    const int w0 = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = L.chan_off[C + 1] - L.chan_off[C];
    for (int w = w0; w < n; w += kSmallWaves) visc_row<MODE, C>(L, L.chan_off[C] + w, lane, X, B, Y);
    __syncthreads();
}

template <class V>
__device__ __forceinline__ void small_rbgs(const ViscLevelView& L, V* X, const V* B, bool xyz) {
    if (xyz) {
        small_pass<kVRedGS, 0, V>(L, X, B, X);
        small_pass<kVBlackGS, 0, V>(L, X, B, X);
        small_pass<kVRedGS, 1, V>(L, X, B, X);
        small_pass<kVBlackGS, 1, V>(L, X, B, X);
        small_pass<kVRedGS, 2, V>(L, X, B, X);
        small_pass<kVBlackGS, 2, V>(L, X, B, X);
    } else {
        small_pass<kVBlackGS, 2, V>(L, X, B, X);
        small_pass<kVRedGS, 2, V>(L, X, B, X);
        small_pass<kVBlackGS, 1, V>(L, X, B, X);
        small_pass<kVRedGS, 1, V>(L, X, B, X);
        small_pass<kVBlackGS, 0, V>(L, X, B, X);
        small_pass<kVRedGS, 0, V>(L, X, B, X);
    }
}

// (zero X when precon,) 2 XYZ sweeps, residual into T, restriction into Cb
template <class V>
__global__ __launch_bounds__(kSmallThreads) void visc_small_pre_kernel(ViscLevelView L, ViscLevelView Cl, V* X,
                                                                       const V* B, V* T, V* Cb, int precon) {
    if (precon) {
        const size_t n16 = size_t(L.nleaf_dof) * 3 * 512 * sizeof(V) / 16;
        uint4* x16 = reinterpret_cast<uint4*>(X);
        for (size_t i = threadIdx.x; i < n16; i += kSmallThreads) x16[i] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
    }
    small_rbgs(L, X, B, true);
    small_rbgs(L, X, B, true);
    small_pass<kVResidual, 0, V>(L, X, B, T);
    small_pass<kVResidual, 1, V>(L, X, B, T);
    small_pass<kVResidual, 2, V>(L, X, B, T);
    const int w0 = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int w = w0; w < Cl.nleaf_dof * 3; w += kSmallWaves) visc_restrict_row(Cl, L, w / 3, w % 3, lane, T, Cb);
}

// X += P Xc, then 2 ZYX sweeps
template <class V>
__global__ __launch_bounds__(kSmallThreads) void visc_small_post_kernel(ViscLevelView L, V* X, const V* B,
                                                                        const V* Xc) {
    const int w0 = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int w = w0; w < L.nleaf_dof * 3; w += kSmallWaves) visc_prolong_row(L, w / 3, w % 3, lane, Xc, X);
    __syncthreads();
    small_rbgs(L, X, B, false);
    small_rbgs(L, X, B, false);
}

// restriction: coarse (leaf, channel) waves
template <class VF, class VC>
__global__ __launch_bounds__(kVThreads) void visc_restrict_kernel(ViscLevelView Cl, ViscLevelView F, const VF* fine,
                                                                  VC* coarse) {
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    if (w >= Cl.nleaf_dof * 3) return;
    visc_restrict_row(Cl, F, w / 3, w - 3 * (w / 3), threadIdx.x & 63, fine, coarse);
}

// prolongation<true>: fine (leaf, channel) waves add the parent coarse value
template <class VF, class VC>
__global__ __launch_bounds__(kVThreads) void visc_prolong_kernel(ViscLevelView F, const VC* coarse, VF* fine) {
    const int w = dev::logical_block() * kVWaves + (threadIdx.x >> 6);
    if (w >= F.nleaf_dof * 3) return;
    visc_prolong_row(F, w / 3, w - 3 * (w / 3), threadIdx.x & 63, coarse, fine);
}

__global__ __launch_bounds__(kVThreads) void visc_zero_kernel(uint4* v, size_t n16) {
    const size_t i = size_t(blockIdx.x) * kVThreads + threadIdx.x;
    if (i < n16) v[i] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------- coarsest solve
// Eigen::ConjugateGradient<RowMajor, Upper|Lower> with the diagonal
// preconditioner, zero guess, 10 iterations, tolerance = float epsilon
// (construct_exact_solver / solve_exact, simd_viscosity3d.cpp:2598-2657).
// Rows are summed in column order as Eigen's row-major product does; the
// dot-product order of Eigen's vectorised reductions is not reproduced.
constexpr int kCgT = 512;
constexpr int kCgPer = kViscMaxCoarsest / kCgT;

__device__ __forceinline__ float wsum(float v) { return dev::wave_sum_dpp(v); }

__device__ __forceinline__ float2 bsum2(float a, float b, float (*red)[kCgT / 64][2], int& par) {
    a = wsum(a);
    b = wsum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[par][w][0] = a;
        red[par][w][1] = b;
    }
    __syncthreads();
    float2 t = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < kCgT / 64; ++i) {
        t.x += red[par][i][0];
        t.y += red[par][i][1];
    }
    par ^= 1;
    return t;
}

// Two barriers per iteration: the reduction of <p, Ap>, and the reduction of
// ||r||^2 and <r, z>, whose barrier also publishes the new z; the new
// direction's product is formed as A p = A z + beta A p_old (the same vector
// as Eigen's A p, rounded differently), so no third barrier publishes p.
// views != null: a batch of systems, workgroup i solving views[i] (one launch
// for the problems of a merged batch whose coarsest level is the same)
template <class V>
__global__ __launch_bounds__(kCgT) void visc_coarsest_kernel(ViscCoarsestView C0, const ViscCoarsestView* views,
                                                             const V* B, V* X) {
    const ViscCoarsestView C = views ? views[blockIdx.x] : C0;
    __shared__ float sz[kViscMaxCoarsest];
    __shared__ float sq[kViscMaxCoarsest];   // A z of the thread's own rows
    __shared__ float red[2][kCgT / 64][2];
    int par = 0;
    const int n = C.n, tid = threadIdx.x;
    float x[kCgPer], r[kCgPer], z[kCgPer], p[kCgPer], q[kCgPer], inv[kCgPer];
    float rr = 0.f, rz = 0.f;
#pragma unroll
    for (int j = 0; j < kCgPer; ++j) {
        const int d = tid + j * kCgT;
        x[j] = r[j] = z[j] = p[j] = q[j] = inv[j] = 0.f;
        if (d < n) {
            r[j] = ld1v(B + C.pos[d]);
            inv[j] = C.invdiag[d];
            rr += r[j] * r[j];
            z[j] = inv[j] * r[j];
            sz[d] = z[j];
            rz += r[j] * z[j];
        }
    }
    const float2 s0 = bsum2(rr, rz, red, par);   // its barrier publishes z
    const float rhs2 = s0.x;
    const float thr = fmaxf(C.tol * C.tol * rhs2, FLT_MIN);
    if (rhs2 != 0.f && !(rhs2 < thr)) {
        float abs_new = s0.y, beta = 0.f;
        for (int it = 0; it < C.max_iters; ++it) {
            // A z row by row: a row's entries are loaded together (padding is
            // in bounds) and summed in column order up to the row's length;
            // the rows loop is not unrolled (registers), so the sums go
            // through LDS, each read back by the thread that wrote it
#pragma unroll 1
            for (int j = 0; j < kCgPer; ++j) {
                const int d = tid + j * kCgT;
                if (d >= n) break;
                const int len = C.elen[d];
                float t = 0.f;
#pragma unroll
                for (int k = 0; k < kViscMaxRow; ++k) {
                    const size_t at = size_t(k) * size_t(n) + size_t(d);
                    const float m = C.eval[at] * sz[C.ecol[at]];
                    const float u = t + m;
                    t = k < len ? u : t;
                }
                sq[d] = t;
            }
            // p = z + beta p and A p = A z + beta A p (beta = 0 on the first step)
            float pq = 0.f;
#pragma unroll
            for (int j = 0; j < kCgPer; ++j) {
                const int d = tid + j * kCgT;
                if (d < n) {
                    p[j] = z[j] + beta * p[j];
                    q[j] = sq[d] + beta * q[j];
                    pq += p[j] * q[j];
                }
            }
            // the barrier also orders this step's reads of z before the writes below
            const float alpha = abs_new / bsum2(pq, 0.f, red, par).x;
            rr = 0.f;
            rz = 0.f;
#pragma unroll
            for (int j = 0; j < kCgPer; ++j) {
                const int d = tid + j * kCgT;
                if (d < n) {
                    x[j] = x[j] + alpha * p[j];
                    r[j] = r[j] - alpha * q[j];
                    rr += r[j] * r[j];
                    z[j] = inv[j] * r[j];
                    sz[d] = z[j];
                    rz += r[j] * z[j];
                }
            }
            const float2 sums = bsum2(rr, rz, red, par);
            if (sums.x < thr) break;
            const float abs_old = abs_new;
            abs_new = sums.y;
            beta = abs_new / abs_old;
        }
    }
#pragma unroll
    for (int j = 0; j < kCgPer; ++j) {
        const int d = tid + j * kCgT;
        if (d < n) st1v(X + C.pos[d], x[j]);
    }
}

// Systems of at most kCgSmallN rows keep their ELL rows in registers (two
// rows per thread), so an iteration reads only the shared z: the same
// products and sums, in the same order, as visc_coarsest_kernel.
constexpr int kCgSmallN = 1024;

template <class V>
__global__ __launch_bounds__(kCgT) void visc_coarsest_small_kernel(ViscCoarsestView C0, const ViscCoarsestView* views,
                                                                   const V* B, V* X) {
    const ViscCoarsestView C = views ? views[blockIdx.x] : C0;
    constexpr int kPer = kCgSmallN / kCgT;
    __shared__ float sz[kCgSmallN];
    __shared__ float red[2][kCgT / 64][2];
    int par = 0;
    const int n = C.n, tid = threadIdx.x;
    float x[kPer], r[kPer], z[kPer], p[kPer], q[kPer], inv[kPer];
    float ev[kPer][kViscMaxRow];
    int ec[kPer][kViscMaxRow], len[kPer];
    float rr = 0.f, rz = 0.f;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int d = tid + j * kCgT;
        x[j] = r[j] = z[j] = p[j] = q[j] = inv[j] = 0.f;
        len[j] = 0;
#pragma unroll
        for (int k = 0; k < kViscMaxRow; ++k) {
            ev[j][k] = 0.f;
            ec[j][k] = 0;
        }
        if (d < n) {
            len[j] = C.elen[d];
#pragma unroll
            for (int k = 0; k < kViscMaxRow; ++k) {
                const size_t at = size_t(k) * size_t(n) + size_t(d);
                ev[j][k] = C.eval[at];
                ec[j][k] = C.ecol[at];
            }
            r[j] = ld1v(B + C.pos[d]);
            inv[j] = C.invdiag[d];
            rr += r[j] * r[j];
            z[j] = inv[j] * r[j];
            sz[d] = z[j];
            rz += r[j] * z[j];
        }
    }
    const float2 s0 = bsum2(rr, rz, red, par);   // its barrier publishes z
    const float rhs2 = s0.x;
    const float thr = fmaxf(C.tol * C.tol * rhs2, FLT_MIN);
    if (rhs2 != 0.f && !(rhs2 < thr)) {
        float abs_new = s0.y, beta = 0.f;
        for (int it = 0; it < C.max_iters; ++it) {
            float pq = 0.f;
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int d = tid + j * kCgT;
                if (d < n) {
                    float t = 0.f;
#pragma unroll
                    for (int k = 0; k < kViscMaxRow; ++k) {
                        const float m = ev[j][k] * sz[ec[j][k]];
                        const float u = t + m;
                        t = k < len[j] ? u : t;
                    }
                    p[j] = z[j] + beta * p[j];
                    q[j] = t + beta * q[j];
                    pq += p[j] * q[j];
                }
            }
            // the barrier also orders this step's reads of z before the writes below
            const float alpha = abs_new / bsum2(pq, 0.f, red, par).x;
            rr = 0.f;
            rz = 0.f;
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int d = tid + j * kCgT;
                if (d < n) {
                    x[j] = x[j] + alpha * p[j];
                    r[j] = r[j] - alpha * q[j];
                    rr += r[j] * r[j];
                    z[j] = inv[j] * r[j];
                    sz[d] = z[j];
                    rz += r[j] * z[j];
                }
            }
            const float2 sums = bsum2(rr, rz, red, par);
            if (sums.x < thr) break;
            const float abs_old = abs_new;
            abs_new = sums.y;
            beta = abs_new / abs_old;
        }
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int d = tid + j * kCgT;
        if (d < n) st1v(X + C.pos[d], x[j]);
    }
}

template <int MODE, class V>
hipError_t apply_mode(const ViscLevelView& L, int c, const V* x, const V* b, V* y, hipStream_t s) {
    const int n = L.chan_off[c + 1] - L.chan_off[c];
    if (n <= 0) return hipSuccess;
    if (L.half == 2) {
        const unsigned gh = unsigned((4 * n + kVWaves - 1) / kVWaves);
        switch (c) {
            case 0: visc_apply_part_kernel<MODE, 0, 2, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
            case 1: visc_apply_part_kernel<MODE, 1, 2, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
            default: visc_apply_part_kernel<MODE, 2, 2, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
        }
        return hipGetLastError();
    }
    if (L.half) {
        const unsigned gh = unsigned((2 * n + kVWaves - 1) / kVWaves);
        switch (c) {
            case 0: visc_apply_part_kernel<MODE, 0, 4, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
            case 1: visc_apply_part_kernel<MODE, 1, 4, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
            default: visc_apply_part_kernel<MODE, 2, 4, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
        }
        return hipGetLastError();
    }
    const unsigned g = unsigned((n + kVWaves - 1) / kVWaves);
    if (L.wide) {
        switch (c) {
            case 0: visc_apply_wide_kernel<MODE, 0, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
            case 1: visc_apply_wide_kernel<MODE, 1, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
            default: visc_apply_wide_kernel<MODE, 2, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        }
        return hipGetLastError();
    }
    switch (c) {
        case 0: visc_apply_kernel<MODE, 0, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        case 1: visc_apply_kernel<MODE, 1, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        default: visc_apply_kernel<MODE, 2, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
    }
    return hipGetLastError();
}

template <class V>
hipError_t apply_one(const ViscLevelView& L, int mode, int c, const V* x, const V* b, V* y, hipStream_t s) {
    if (c < 0 || c > 2) return hipErrorInvalidValue;
    switch (mode) {
        case kVNormal: return apply_mode<kVNormal>(L, c, x, b, y, s);
        case kVResidual: return apply_mode<kVResidual>(L, c, x, b, y, s);
        case kVJacobi: return apply_mode<kVJacobi>(L, c, x, b, y, s);
        case kVSpai0Mode: return apply_mode<kVSpai0Mode>(L, c, x, b, y, s);
        case kVRedGS: return apply_mode<kVRedGS>(L, c, x, b, y, s);
        case kVBlackGS: return apply_mode<kVBlackGS>(L, c, x, b, y, s);
        default: return hipErrorInvalidValue;
    }
}

template <class V>
hipError_t apply_all(const ViscLevelView& L, int mode, const V* x, const V* b, V* y, hipStream_t s) {
    const int n = L.chan_off[3];
    if (n <= 0) return hipSuccess;
    if (L.half) {
        // the three channels' rows in one launch on half rows (quarter rows too only for the passes)
        const unsigned gh = unsigned((2 * n + kVWaves - 1) / kVWaves);
        switch (mode) {
            case kVNormal: visc_apply_all_part_kernel<kVNormal, 4, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
            case kVResidual: visc_apply_all_part_kernel<kVResidual, 4, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
            case kVJacobi: visc_apply_all_part_kernel<kVJacobi, 4, V><<<gh, kVThreads, 0, s>>>(L, x, b, y); break;
            case kVSpai0Mode:
                visc_apply_all_part_kernel<kVSpai0Mode, 4, V><<<gh, kVThreads, 0, s>>>(L, x, b, y);
                break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const unsigned g = unsigned((n + kVWaves - 1) / kVWaves);
    if (L.wide) {
        switch (mode) {
            case kVNormal: visc_apply_all_wide_kernel<kVNormal, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
            case kVResidual: visc_apply_all_wide_kernel<kVResidual, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
            case kVJacobi: visc_apply_all_wide_kernel<kVJacobi, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
            case kVSpai0Mode: visc_apply_all_wide_kernel<kVSpai0Mode, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (mode) {
        case kVNormal: visc_apply_all_kernel<kVNormal, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        case kVResidual: visc_apply_all_kernel<kVResidual, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        case kVJacobi: visc_apply_all_kernel<kVJacobi, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        case kVSpai0Mode: visc_apply_all_kernel<kVSpai0Mode, V><<<g, kVThreads, 0, s>>>(L, x, b, y); break;
        default: return hipErrorInvalidValue;   // Gauss-Seidel passes update in place, channel by channel
    }
    return hipGetLastError();
}

template <class VF, class VC>
hipError_t restrict_to(const ViscLevelView& coarse, const ViscLevelView& fine, const VF* fine_v, VC* coarse_v,
                       hipStream_t s) {
    const int n = coarse.nleaf_dof * 3;
    if (n <= 0) return hipSuccess;
    visc_restrict_kernel<VF, VC><<<unsigned((n + kVWaves - 1) / kVWaves), kVThreads, 0, s>>>(coarse, fine, fine_v,
                                                                                             coarse_v);
    return hipGetLastError();
}

template <class VF, class VC>
hipError_t prolong_from(const ViscLevelView& fine, const VC* coarse_v, VF* fine_v, hipStream_t s) {
    const int n = fine.nleaf_dof * 3;
    if (n <= 0) return hipSuccess;
    visc_prolong_kernel<VF, VC><<<unsigned((n + kVWaves - 1) / kVWaves), kVThreads, 0, s>>>(fine, coarse_v, fine_v);
    return hipGetLastError();
}

template <class V>
hipError_t coarsest(const ViscCoarsestView& C, const V* b, V* x, hipStream_t s) {
    if (C.n <= 0) return hipSuccess;
    if (C.n > kViscMaxCoarsest) return hipErrorInvalidValue;
    if (C.n <= kCgSmallN) visc_coarsest_small_kernel<V><<<1, kCgT, 0, s>>>(C, nullptr, b, x);
    else visc_coarsest_kernel<V><<<1, kCgT, 0, s>>>(C, nullptr, b, x);
    return hipGetLastError();
}

template <class V>
hipError_t bcoarsest(const ViscCoarsestView* views, int count, int max_n, const V* b, V* x, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (max_n > kViscMaxCoarsest) return hipErrorInvalidValue;
    // the register form when every system fits it, as each problem alone would run
    if (max_n <= kCgSmallN) visc_coarsest_small_kernel<V><<<unsigned(count), kCgT, 0, s>>>({}, views, b, x);
    else visc_coarsest_kernel<V><<<unsigned(count), kCgT, 0, s>>>({}, views, b, x);
    return hipGetLastError();
}

template <class V>
hipError_t zero(const ViscLevelView& L, V* v, hipStream_t s) {
    const size_t n16 = size_t(L.nleaf_dof) * 3 * 512 * sizeof(V) / 16;
    if (!n16) return hipSuccess;
    visc_zero_kernel<<<unsigned((n16 + kVThreads - 1) / kVThreads), kVThreads, 0, s>>>(reinterpret_cast<uint4*>(v),
                                                                                      n16);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_visc_apply(const ViscLevelView& L, int mode, int c, const float* x, const float* b, float* y,
                             hipStream_t s) {
    return apply_one(L, mode, c, x, b, y, s);
}
hipError_t launch_visc_apply(const ViscLevelView& L, int mode, int c, const bf16* x, const bf16* b, bf16* y,
                             hipStream_t s) {
    return apply_one(L, mode, c, x, b, y, s);
}

hipError_t launch_visc_apply_all(const ViscLevelView& L, int mode, const float* x, const float* b, float* y,
                                 hipStream_t s) {
    return apply_all(L, mode, x, b, y, s);
}
hipError_t launch_visc_apply_all(const ViscLevelView& L, int mode, const bf16* x, const bf16* b, bf16* y,
                                 hipStream_t s) {
    return apply_all(L, mode, x, b, y, s);
}

hipError_t launch_visc_restrict(const ViscLevelView& coarse, const ViscLevelView& fine, const float* fine_v,
                                float* coarse_v, hipStream_t s) {
    return restrict_to(coarse, fine, fine_v, coarse_v, s);
}
hipError_t launch_visc_restrict(const ViscLevelView& coarse, const ViscLevelView& fine, const float* fine_v,
                                bf16* coarse_v, hipStream_t s) {
    return restrict_to(coarse, fine, fine_v, coarse_v, s);
}
hipError_t launch_visc_restrict(const ViscLevelView& coarse, const ViscLevelView& fine, const bf16* fine_v,
                                bf16* coarse_v, hipStream_t s) {
    return restrict_to(coarse, fine, fine_v, coarse_v, s);
}

hipError_t launch_visc_prolong(const ViscLevelView& fine, const ViscLevelView&, const float* coarse_v, float* fine_v,
                               hipStream_t s) {
    return prolong_from(fine, coarse_v, fine_v, s);
}
hipError_t launch_visc_prolong(const ViscLevelView& fine, const ViscLevelView&, const bf16* coarse_v, float* fine_v,
                               hipStream_t s) {
    return prolong_from(fine, coarse_v, fine_v, s);
}
hipError_t launch_visc_prolong(const ViscLevelView& fine, const ViscLevelView&, const bf16* coarse_v, bf16* fine_v,
                               hipStream_t s) {
    return prolong_from(fine, coarse_v, fine_v, s);
}

hipError_t launch_visc_coarsest(const ViscCoarsestView& C, const float* b, float* x, hipStream_t s) {
    return coarsest(C, b, x, s);
}
hipError_t launch_visc_coarsest(const ViscCoarsestView& C, const bf16* b, bf16* x, hipStream_t s) {
    return coarsest(C, b, x, s);
}
hipError_t launch_visc_bcoarsest(const ViscCoarsestView* views, int count, int max_n, const float* b, float* x,
                                 hipStream_t s) {
    return bcoarsest(views, count, max_n, b, x, s);
}

hipError_t launch_visc_small_pre(const ViscLevelView& L, const ViscLevelView& coarse, float* x, const float* b,
                                 float* t, float* coarse_b, bool precon, hipStream_t s) {
    visc_small_pre_kernel<float><<<1, kSmallThreads, 0, s>>>(L, coarse, x, b, t, coarse_b, precon ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_visc_small_pre(const ViscLevelView& L, const ViscLevelView& coarse, bf16* x, const bf16* b,
                                 bf16* t, bf16* coarse_b, bool precon, hipStream_t s) {
    visc_small_pre_kernel<bf16><<<1, kSmallThreads, 0, s>>>(L, coarse, x, b, t, coarse_b, precon ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_visc_small_post(const ViscLevelView& L, float* x, const float* b, const float* xc, hipStream_t s) {
    visc_small_post_kernel<float><<<1, kSmallThreads, 0, s>>>(L, x, b, xc);
    return hipGetLastError();
}
hipError_t launch_visc_small_post(const ViscLevelView& L, bf16* x, const bf16* b, const bf16* xc, hipStream_t s) {
    visc_small_post_kernel<bf16><<<1, kSmallThreads, 0, s>>>(L, x, b, xc);
    return hipGetLastError();
}

hipError_t launch_visc_zero(const ViscLevelView& L, float* v, hipStream_t s) { return zero(L, v, s); }
hipError_t launch_visc_zero(const ViscLevelView& L, bf16* v, hipStream_t s) { return zero(L, v, s); }

}  // namespace uaamg
