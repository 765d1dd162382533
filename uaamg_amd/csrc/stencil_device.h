// Device-side building blocks shared by the level kernels and the PCG kernels:
// row loads, the 7-point stencil over a pluggable row source, one Jacobi row.
//
// Arithmetic follows LaplacianApplySIMD::LightWeightApplier
// (simd_vdb_poisson.cpp:755-1031) term by term: the six neighbour products are
// rounded separately and summed as ((x+ + x-) + (y+ + y-)) + (z+ + z-), then
// A*x = fma(x, diag, sum). Sources are compiled with -ffp-contract=off so the
// only fused multiply-adds are the explicit fmaf calls.
#pragma once

#include "poisson_kernels.h"
#include "xcd_map.h"

namespace uaamg {
namespace dev {

__device__ __forceinline__ int wave_leaf() {
    int leaf = logical_block() * kWavesPerBlock + int(threadIdx.x >> 6);
    return __builtin_amdgcn_readfirstlane(leaf);
}

__device__ __forceinline__ void ld8(const float* p, float v[8]) {
    const float4 a = reinterpret_cast<const float4*>(p)[0];
    const float4 b = reinterpret_cast<const float4*>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

__device__ __forceinline__ void st8(float* p, const float v[8]) {
    reinterpret_cast<float4*>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<float4*>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}

// BF16 storage (W-cycle vectors of the mixed-precision form): a z-row is 16
// bytes; widening is exact, narrowing rounds to nearest even
// (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ void ld8(const bf16* p, float v[8]) {
    const uint4 q = *reinterpret_cast<const uint4*>(p);
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[2 * i] = __uint_as_float(w[i] << 16);
        v[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
    }
}

__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    const bf16 ha = bf16(a), hb = bf16(b);
    return uint32_t(__builtin_bit_cast(uint16_t, ha)) | (uint32_t(__builtin_bit_cast(uint16_t, hb)) << 16);
}

__device__ __forceinline__ void st8(bf16* p, const float v[8]) {
    *reinterpret_cast<uint4*>(p) =
        make_uint4(pack_bf16(v[0], v[1]), pack_bf16(v[2], v[3]), pack_bf16(v[4], v[5]), pack_bf16(v[6], v[7]));
}

__device__ __forceinline__ float ld1(const float* p) { return *p; }
__device__ __forceinline__ float ld1(const bf16* p) { return float(*p); }
__device__ __forceinline__ void st1(float* p, float v) { *p = v; }
__device__ __forceinline__ void st1(bf16* p, float v) { *p = bf16(v); }

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ float4 ld4(const bf16* p) {
    const uint2 q = *reinterpret_cast<const uint2*>(p);
    return make_float4(__uint_as_float(q.x << 16), __uint_as_float(q.x & 0xffff0000u), __uint_as_float(q.y << 16),
                       __uint_as_float(q.y & 0xffff0000u));
}

__device__ __forceinline__ void fill8(float v[8], float c) {
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = c;
}

__device__ __forceinline__ uint32_t row_mask(const LevelView& L, int leaf, int lane) {
    const uint64_t w = L.mask[size_t(leaf) * 8 + (lane >> 3)];
    return uint32_t(w >> ((lane & 7) * 8)) & 0xffu;
}

// load a row of a brick array, or a constant when the slot is absent
__device__ __forceinline__ void ld8_slot(const float* base, int slot, int row, float dflt, float v[8]) {
    if (slot >= 0) ld8(base + size_t(slot) * 512 + row * 8, v);
    else fill8(v, dflt);
}

// A leaf's record (slots, neighbour slots, face neighbours, parents; layout
// at kSlotDiag), loaded once per wavefront at the top of a kernel.
struct LeafRec {
    int sd, sx, sy, sz, ssp, sxp, syp, szp;
    int nb[6];
    int par;
    int nbsd[6], nbpar[6];
};

__device__ __forceinline__ LeafRec leaf_rec(const LevelView& L, int leaf) {
    const int4* p = reinterpret_cast<const int4*>(L.slot + size_t(leaf) * kSlotStride);
    const int4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5], g = p[6];
    LeafRec r;
    r.sd = a.x; r.sx = a.y; r.sy = a.z; r.sz = a.w;
    r.ssp = b.x; r.sxp = b.y; r.syp = b.z; r.szp = b.w;
    r.nb[0] = c.x; r.nb[1] = c.y; r.nb[2] = c.z; r.nb[3] = c.w;
    r.nb[4] = d.x; r.nb[5] = d.y; r.par = d.z;
    r.nbsd[0] = e.x; r.nbsd[1] = e.y; r.nbsd[2] = e.z; r.nbsd[3] = e.w;
    r.nbsd[4] = f.x; r.nbsd[5] = f.y; r.nbpar[0] = f.z; r.nbpar[1] = f.w;
    r.nbpar[2] = g.x; r.nbpar[3] = g.y; r.nbpar[4] = g.z; r.nbpar[5] = g.w;
    // One use of every field here, in the kernel's first basic block: the
    // record's scalar loads are then issued together and waited for once
    // (otherwise each is issued in the branch that first reads it, one
    // memory round trip after the other). The empty asm emits no instruction.
    const int all = (((a.x ^ a.y) ^ (a.z ^ a.w)) ^ ((b.x ^ b.y) ^ (b.z ^ b.w))) ^
                    (((c.x ^ c.y) ^ (c.z ^ c.w)) ^ ((d.x ^ d.y) ^ (d.z ^ d.w))) ^
                    (((e.x ^ e.y) ^ (e.z ^ e.w)) ^ ((f.x ^ f.y) ^ (f.z ^ f.w))) ^ ((g.x ^ g.y) ^ (g.z ^ g.w));
    asm volatile("" ::"s"(all));
    return r;
}

// Sources of the vector a stencil is applied to. row() fills the 8 values of
// one z-row of a leaf, at() returns one value; both are called only for leaves
// that exist (an absent neighbour leaf contributes 0), with that leaf's
// diagonal slot sd and parent par from the leaf records.
template <class V>
struct PlainSrcT {   // the stored brick vector
    const V* X;
    __device__ __forceinline__ void row(int leaf, int, int, int r, float v[8]) const {
        ld8(X + size_t(leaf) * 512 + r * 8, v);
    }
    __device__ __forceinline__ float at(int leaf, int, int, int r, int z) const {
        return ld1(X + size_t(leaf) * 512 + r * 8 + z);
    }
};
using PlainSrc = PlainSrcT<float>;

// x1 = w0 * D^-1 b: the first Jacobi sweep from a zero guess, evaluated where
// the second sweep needs it (setGridToResultAfterFirstJacobi,
// simd_vdb_poisson.cpp:1435-1479: (invdiag * rhs) * w). B is 0 on non-DOF
// voxels, so x1 is too.
template <class VB>
struct FirstJacobiSrcT {
    const VB* B;
    const float* invdiag;
    float dflt_inv;
    float w0;
    __device__ __forceinline__ void row(int leaf, int sd, int, int r, float v[8]) const {
        float b[8], inv[8];
        ld8(B + size_t(leaf) * 512 + r * 8, b);
        ld8_slot(invdiag, sd, r, dflt_inv, inv);
#pragma unroll
        for (int z = 0; z < 8; ++z) v[z] = w0 * (inv[z] * b[z]);
    }
    __device__ __forceinline__ float at(int leaf, int sd, int, int r, int z) const {
        const float inv = sd >= 0 ? invdiag[size_t(sd) * 512 + r * 8 + z] : dflt_inv;
        return w0 * (inv * ld1(B + size_t(leaf) * 512 + r * 8 + z));
    }
};
using FirstJacobiSrc = FirstJacobiSrcT<float>;

// x + P x_coarse on on-voxels (prolongation<true> with alpha = 1,
// simd_vdb_poisson.cpp:1686-1754), evaluated where the following sweep needs
// it. The coarse row is loaded whenever the leaf has a parent and added on
// the row's DOF voxels only.
template <class V>
struct ProlongSrcT {
    const V* X;
    const V* Xc;
    const uint64_t* mask;
    __device__ __forceinline__ uint32_t bits(int leaf, int r) const {
        return uint32_t(mask[size_t(leaf) * 8 + (r >> 3)] >> ((r & 7) * 8)) & 0xffu;
    }
    __device__ __forceinline__ float4 coarse4(int par, int r) const {
        const int cl = par >> 3, oct = par & 7;
        const int crow = (((r >> 3) >> 1) + 4 * ((oct >> 2) & 1)) * 8 + (((r & 7) >> 1) + 4 * ((oct >> 1) & 1));
        return ld4(Xc + size_t(cl) * 512 + crow * 8 + 4 * (oct & 1));
    }
    __device__ __forceinline__ void row(int leaf, int, int par, int r, float v[8]) const {
        ld8(X + size_t(leaf) * 512 + r * 8, v);
        if (par >= 0) {   // bounds guard: every DOF leaf has a parent
            const uint32_t m = bits(leaf, r);
            const float4 c = coarse4(par, r);
            const float cc[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
            for (int z = 0; z < 8; ++z)
                if ((m >> z) & 1u) v[z] += cc[z >> 1];
        }
    }
    __device__ __forceinline__ float at(int leaf, int, int par, int r, int z) const {
        float v = ld1(X + size_t(leaf) * 512 + r * 8 + z);
        if (par >= 0) {
            const uint32_t m = bits(leaf, r);
            const float4 c = coarse4(par, r);
            const float cc[4] = {c.x, c.y, c.z, c.w};
            if ((m >> z) & 1u) v += cc[z >> 1];
        }
        return v;
    }
};
using ProlongSrc = ProlongSrcT<float>;

// p' = z + beta p (levelXPlusAlphaY, simd_vdb_poisson.cpp:2305-2320), both 0 off the DOFs
struct UpdateSrc {
    const float* Z;
    const float* P;
    float beta;
    __device__ __forceinline__ void row(int leaf, int, int, int r, float v[8]) const {
        float z[8], p[8];
        ld8(Z + size_t(leaf) * 512 + r * 8, z);
        ld8(P + size_t(leaf) * 512 + r * 8, p);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float t = beta * p[k];
            v[k] = z[k] + t;
        }
    }
    __device__ __forceinline__ float at(int leaf, int, int, int r, int z) const {
        const size_t i = size_t(leaf) * 512 + r * 8 + z;
        const float t = beta * P[i];
        return Z[i] + t;
    }
};

// The PCG iterate with its pending update, x + alpha p, formed as the update
// kernels form it (product, then sum; contraction is off), so that values
// read here equal the ones the deferred update later stores.
struct StepSrc {
    const float* X;
    const float* P;
    float alpha;
    __device__ __forceinline__ void row(int leaf, int, int, int r, float v[8]) const {
        float x[8], p[8];
        ld8(X + size_t(leaf) * 512 + r * 8, x);
        ld8(P + size_t(leaf) * 512 + r * 8, p);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float t = alpha * p[k];
            v[k] = x[k] + t;
        }
    }
    __device__ __forceinline__ float at(int leaf, int, int, int r, int z) const {
        const size_t i = size_t(leaf) * 512 + r * 8 + z;
        const float t = alpha * P[i];
        return X[i] + t;
    }
};

// The operands of the 7-point stencil for one row: neighbour values and the
// coefficients of the faces they couple through.
struct StencilRow {
    float xm[8], xp[8], ym[8], yp[8];
    float zlo, zhi;
    float cxm[8], cxp[8], cym[8], cyp[8], czm[8];
    float czhi;
};

// Own row of the source and the off-diagonal operands of one row.
template <class Src>
__device__ __forceinline__ void stencil_gather(const LevelView& L, const LeafRec& R, const Src& src, int leaf, int lane,
                                               float xc[8], StencilRow& o) {
    const int lx = lane >> 3, ly = lane & 7;
    const int nmz = R.nb[4], npz = R.nb[5];
    src.row(leaf, R.sd, R.par, lane, xc);

    float* xm = o.xm;
    float* xp = o.xp;
    float* ym = o.ym;
    float* yp = o.yp;
    // rows of the own leaf or, on the brick face, of the face neighbour f
    auto side = [&](bool own, int f, int rown, int rnb, float v[8]) {
        const int sl = own ? leaf : R.nb[f];
        const int ssd = own ? R.sd : R.nbsd[f];
        const int spar = own ? R.par : R.nbpar[f];
        if (sl >= 0) src.row(sl, ssd, spar, own ? rown : rnb, v);
        else fill8(v, 0.f);
    };
    side(lx > 0, 0, lane - 8, lane + 56, xm);
    side(lx < 7, 1, lane + 8, lane - 56, xp);
    side(ly > 0, 2, lane - 1, lane + 7, ym);
    side(ly < 7, 3, lane + 1, lane - 7, yp);
    o.zlo = nmz >= 0 ? src.at(nmz, R.nbsd[4], R.nbpar[4], lane, 7) : 0.f;
    o.zhi = npz >= 0 ? src.at(npz, R.nbsd[5], R.nbpar[5], lane, 0) : 0.f;

    const int sx = R.sx, sy = R.sy, sz = R.sz, sxp = R.sxp, syp = R.syp, szp = R.szp;
    const float d = L.dflt_off;

    ld8_slot(L.cx, sx, lane, d, o.cxm);
    ld8_slot(L.cx, lx < 7 ? sx : sxp, lx < 7 ? lane + 8 : lane - 56, d, o.cxp);
    ld8_slot(L.cy, sy, lane, d, o.cym);
    ld8_slot(L.cy, ly < 7 ? sy : syp, ly < 7 ? lane + 1 : lane - 7, d, o.cyp);
    ld8_slot(L.cz, sz, lane, d, o.czm);
    o.czhi = szp >= 0 ? L.cz[size_t(szp) * 512 + lane * 8] : d;
}

// Own row of the source and the off-diagonal sum of the 7-point stencil for one row.
template <class Src>
__device__ __forceinline__ void stencil_offdiag(const LevelView& L, const LeafRec& R, const Src& src, int leaf,
                                                int lane, float xc[8], float s[8]) {
    StencilRow o;
    stencil_gather(L, R, src, leaf, lane, xc, o);
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const float zpv = z < 7 ? xc[z + 1] : o.zhi;
        const float czp = z < 7 ? o.czm[z + 1] : o.czhi;
        const float zmv = z > 0 ? xc[z - 1] : o.zlo;
        const float f1 = o.xp[z] * o.cxp[z];
        const float f2 = o.xm[z] * o.cxm[z];
        const float f3 = o.yp[z] * o.cyp[z];
        const float f4 = o.ym[z] * o.cym[z];
        const float f5 = zpv * czp;
        const float f6 = zmv * o.czm[z];
        s[z] = ((f1 + f2) + (f3 + f4)) + (f5 + f6);
    }
}

// The same off-diagonal sum with every product and sum in FP64 (exact
// products of FP32 operands), for the true residual of the refinement step.
template <class Src>
__device__ __forceinline__ void stencil_offdiag64(const LevelView& L, const LeafRec& R, const Src& src, int leaf,
                                                  int lane, float xc[8], double s[8]) {
    StencilRow o;
    stencil_gather(L, R, src, leaf, lane, xc, o);
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const double zpv = z < 7 ? xc[z + 1] : o.zhi;
        const double czp = z < 7 ? o.czm[z + 1] : o.czhi;
        const double zmv = z > 0 ? xc[z - 1] : o.zlo;
        const double f1 = double(o.xp[z]) * double(o.cxp[z]);
        const double f2 = double(o.xm[z]) * double(o.cxm[z]);
        const double f3 = double(o.yp[z]) * double(o.cyp[z]);
        const double f4 = double(o.ym[z]) * double(o.cym[z]);
        const double f5 = zpv * czp;
        const double f6 = zmv * double(o.czm[z]);
        s[z] = ((f1 + f2) + (f3 + f4)) + (f5 + f6);
    }
}


// Weighted Jacobi on one row: y = fma(w, (b - (fma(x, diag, s))) * D^-1, x).
template <class VB>
__device__ __forceinline__ void jacobi_row(const LevelView& L, const LeafRec& R, int leaf, int lane, const float xc[8],
                                           const float s[8], const VB* B, float w, float y[8]) {
    const int sd = R.sd;
    float dg[8], inv[8], b[8];
    ld8_slot(L.diag, sd, lane, L.dflt_diag, dg);
    ld8_slot(L.invdiag, sd, lane, L.dflt_invdiag, inv);
    ld8(B + size_t(leaf) * 512 + lane * 8, b);
#pragma unroll
    for (int z = 0; z < 8; ++z) {
        const float ax = fmaf(xc[z], dg[z], s[z]);
        float r = b[z] - ax;
        r = r * inv[z];
        y[z] = fmaf(w, r, xc[z]);
    }
}

__device__ __forceinline__ void mask_row(uint32_t rm, float y[8]) {
#pragma unroll
    for (int z = 0; z < 8; ++z)
        if (!((rm >> z) & 1u)) y[z] = 0.f;
}

// index of the first DOF of a leaf's z-row (DOFs numbered in leaf order, then voxel offset)
__device__ __forceinline__ int64_t row_dof_base(const LevelView& L, int leaf, int lane) {
    const uint64_t* m = L.mask + size_t(leaf) * 8;
    const int lx = lane >> 3, ly = lane & 7;
    int64_t base = L.dofbase[leaf];
    for (int w = 0; w < lx; ++w) base += __popcll(m[w]);
    const uint64_t below = ly == 0 ? 0ull : (m[lx] & ((uint64_t(1) << (ly * 8)) - 1));
    return base + __popcll(below);
}

inline unsigned grid_for(int32_t nleaf) { return unsigned((nleaf + kWavesPerBlock - 1) / kWavesPerBlock); }

}  // namespace dev
}  // namespace uaamg
