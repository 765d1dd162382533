// ORACLE - TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the SIMD_UAAMG pressure-Poisson solve path of the
// reference (SIMD_UAAMG/simd_vdb_poisson.{h,cpp}, SIMD_UAAMG_Ops.h,
// openvdb_grid_math_op.h, vdb_SIMD_IO.h). It is the checker for the HIP
// product: only tests, the smoke entry point and the benchmark's CPU-baseline
// leg may load it. Nothing in the product calls it.
//
// Structure: grids are std::map<origin, leaf> kept in OpenVDB Tree4<5,4,3>
// traversal order, read through accessor-style get() calls with the
// reference's background/trimmed-leaf semantics. The apply operator works on
// linearised leaf arrays with a 6-face neighbour table as the reference's
// LaplacianApplySIMD::initLinearizedLeaves does (:1192-1245), evaluated per
// voxel in the same summation order as the AVX code. Build with
// -ffp-contract=off: the only fused multiply-adds are explicit fma() calls.
//
// Every solve routine exists in an FP32 form (reference arithmetic; dot
// products accumulate in float per leaf, leaves summed in tree order) and an
// FP64 twin (same algorithm, double vectors and arithmetic on the same float
// coefficients) used as the "FP64 oracle" solution.
//
// Parity status of each function is pinned by tests/test_oracle_*.py; what no
// test pins is marked "parity unpinned" below.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "vdb_grid.h"

namespace orc {

float fraction_inside(float phi_left, float phi_right) {
    if (phi_left < 0 && phi_right < 0) return 1;
    if (phi_left < 0 && phi_right >= 0) return phi_left / (phi_left - phi_right);
    if (phi_left >= 0 && phi_right < 0) return phi_right / (phi_right - phi_left);
    return 0;
}

// 4-corner area fraction (levelset_util.cpp:22-103)
float fraction_inside4(float bl, float br, float tl, float tr) {
    const int n_in = (bl < 0) + (tl < 0) + (br < 0) + (tr < 0);
    float c[4] = {bl, br, tr, tl};
    auto cycle = [&]() { std::rotate(c, c + 1, c + 4); };
    switch (n_in) {
        case 4: return 1;
        case 3: {
            while (c[0] < 0) cycle();
            const float a = 1 - fraction_inside(c[0], c[3]);
            const float b = 1 - fraction_inside(c[0], c[1]);
            return 1 - 0.5f * a * b;
        }
        case 2: {
            while (c[0] >= 0 || !(c[1] < 0 || c[2] < 0)) cycle();
            if (c[1] < 0) return 0.5f * (fraction_inside(c[0], c[3]) + fraction_inside(c[1], c[2]));
            const float mid = 0.25f * (c[0] + c[1] + c[2] + c[3]);
            if (mid < 0) {
                float area = 0;
                area += 0.5f * (1 - fraction_inside(c[0], c[3])) * (1 - fraction_inside(c[2], c[3]));
                area += 0.5f * (1 - fraction_inside(c[0], c[1])) * (1 - fraction_inside(c[2], c[1]));
                return 1 - area;
            }
            float area = 0;
            area += 0.5f * fraction_inside(c[0], c[1]) * fraction_inside(c[0], c[3]);
            area += 0.5f * fraction_inside(c[2], c[1]) * fraction_inside(c[2], c[3]);
            return area;
        }
        case 1: {
            while (c[0] >= 0) cycle();
            return 0.5f * fraction_inside(c[0], c[3]) * fraction_inside(c[0], c[1]);
        }
        default: return 0;
    }
}

// ------------------------------------------------------------------ level
struct Level {
    int level = 0;
    float dt = 0, dx = 0;
    int64_t ndof = 0;
    IGrid dof;                     // DOF index; keeps empty leaves like the reference tree
    FGrid diag, xe, ye, ze;        // trimmed coefficient grids
    FGrid inv, spai0;
    bool spai0_ready = false;

    // linearised leaves (LeafManager order of the DOF tree)
    std::vector<Coord> org;
    std::vector<const Leaf<int32_t>*> dl;
    std::vector<std::array<int64_t, 6>> nb;   // +x,-x,+y,-y,+z,-z ; -1 = none
    std::vector<const float*> cdiag, cinv, cspai, cx, cy, cz, cxu, cyu, czu;
    float t = 0;   // dt/dx^2
    float d_face = 0, d_diag = 0, d_inv = 0, d_spai = 0;

    size_t nleaf() const { return org.size(); }
};

const float* data_or_null(const FGrid& g, Coord o) {
    const Leaf<float>* l = g.probe(o);
    return l ? l->v : nullptr;
}

void set_dof_index(Level& L) {
    // per-leaf count, prefix scan in leaf order, then running index (:637-671)
    int64_t run = 0;
    for (auto& kv : L.dof.leaves) {
        Leaf<int32_t>& lf = kv.second;
        for (int i = 0; i < 512; ++i) {
            if (lf.is_on(i)) lf.v[i] = int32_t(run++);
            else lf.v[i] = -1;
        }
    }
    L.ndof = run;
}

void trim(FGrid& g, float dflt, float eps) {
    eps = std::fabs(eps);
    std::vector<Coord> drop;
    for (auto& kv : g.leaves) {
        float max_err = 0;
        for (int i = 0; i < 512; ++i)
            if (kv.second.is_on(i)) max_err = std::max(max_err, std::fabs(kv.second.v[i] - dflt));
        if (max_err <= eps) drop.push_back(kv.first);
    }
    for (const Coord& c : drop) g.leaves.erase(c);
    g.background = dflt;   // stolen leaves become tiles of the default value
}

void trim_level(Level& L) {
    const float term = -L.dt / (L.dx * L.dx);
    const float dterm = -6.0f * term;
    trim(L.diag, dterm, dterm * 1e-5f);
    trim(L.xe, term, term * 1e-5f);
    trim(L.ye, term, term * 1e-5f);
    trim(L.ze, term, term * 1e-5f);
}

void linearise(Level& L) {
    L.t = L.dt / (L.dx * L.dx);
    L.d_face = -L.t;
    L.d_diag = 6.f * L.t;
    L.d_inv = 1.0f / (6.f * L.t);
    L.d_spai = 1.0f / (7.f * L.t);
    // inverse diagonal on the trimmed diagonal topology (initInvDiagonal :1081-1111)
    L.inv.topology_copy(L.diag, 1.0f / L.diag.background);
    for (auto& kv : L.inv.leaves) {
        const Leaf<float>* d = L.diag.probe(kv.first);
        for (int i = 0; i < 512; ++i)
            if (kv.second.is_on(i)) kv.second.v[i] = d->v[i] == 0 ? 0.f : 1.0f / d->v[i];
    }
    L.org.clear();
    L.dl.clear();
    std::map<Coord, int64_t, TreeLess> pos;
    for (const auto& kv : L.dof.leaves) {
        pos[kv.first] = int64_t(L.org.size());
        L.org.push_back(kv.first);
        L.dl.push_back(&kv.second);
    }
    const size_t n = L.org.size();
    L.nb.assign(n, {-1, -1, -1, -1, -1, -1});
    L.cdiag.resize(n);
    L.cinv.resize(n);
    L.cx.resize(n);
    L.cy.resize(n);
    L.cz.resize(n);
    L.cxu.resize(n);
    L.cyu.resize(n);
    L.czu.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const Coord o = L.org[i];
        for (int f = 0; f < 6; ++f) {
            Coord q = o;
            q[f / 2] += (f % 2 == 0) ? 8 : -8;
            auto it = pos.find(q);
            if (it != pos.end()) L.nb[i][f] = it->second;
        }
        L.cdiag[i] = data_or_null(L.diag, o);
        L.cinv[i] = data_or_null(L.inv, o);
        L.cx[i] = data_or_null(L.xe, o);
        L.cy[i] = data_or_null(L.ye, o);
        L.cz[i] = data_or_null(L.ze, o);
        L.cxu[i] = data_or_null(L.xe, Coord{o.x + 8, o.y, o.z});
        L.cyu[i] = data_or_null(L.ye, Coord{o.x, o.y + 8, o.z});
        L.czu[i] = data_or_null(L.ze, Coord{o.x, o.y, o.z + 8});
    }
    L.cspai.assign(n, nullptr);
}

// BuildFinestMatrix + initializeFinest (:474-635)
void build_finest(Level& L, const FGrid& phi, const VGrid& w, float dt, float dx) {
    L.level = 0;
    L.dt = dt;
    L.dx = dx;
    const float term = dt / (dx * dx);
    L.diag.topology_copy(phi, 6 * term);
    L.xe.topology_copy(phi, -term);
    L.ye.topology_copy(phi, -term);
    L.ze.topology_copy(phi, -term);
    for (auto& kv : phi.leaves) {
        const Coord o = kv.first;
        const Leaf<float>& pl = kv.second;
        Leaf<float>& dleaf = *L.diag.probe(o);
        Leaf<float>* tl[3] = {L.xe.probe(o), L.ye.probe(o), L.ze.probe(o)};
        for (int off = 0; off < 512; ++off) {
            if (!pl.is_on(off)) continue;
            const float ph = pl.v[off];
            bool dof = false;
            if (ph < 0.f) {
                const Coord lc = local_of(off);
                const Coord g{o.x + lc.x, o.y + lc.y, o.z + lc.z};
                float diagonal = 0.f, txyz[3] = {0.f, 0.f, 0.f};
                for (int f = 0; f < 6; ++f) {
                    const int c = f / 2;
                    const bool pos = f % 2 == 0;
                    Coord wc = g, pc = g;
                    if (pos) {
                        wc[c]++;
                        pc[c]++;
                    } else {
                        pc[c]--;
                    }
                    const float weight = w.get(wc).v[c];
                    const float other = phi.get(pc);
                    const float tt = weight * term;
                    if (other < 0) {
                        diagonal += tt;
                        if (!pos) txyz[c] = -tt;
                    } else {
                        float theta = fraction_inside(ph, other);
                        if (theta < 0.02f) theta = 0.02f;
                        diagonal += tt / theta;
                    }
                }
                if (diagonal != 0.f) {
                    dof = true;
                    dleaf.v[off] = diagonal;
                    for (int c = 0; c < 3; ++c) tl[c]->v[off] = txyz[c];
                }
            }
            dleaf.set_on(off, dof);
            for (int c = 0; c < 3; ++c) tl[c]->set_on(off, dof);
        }
    }
    L.dof.topology_copy(L.diag, -1);
    set_dof_index(L);
    trim_level(L);
    linearise(L);
}

// initializeFromFineLevel (:260-471) with TouchCoarseLeafReducer (SIMD_UAAMG_Ops.h:4-37)
void build_coarse(Level& C, const Level& F) {
    C.level = F.level + 1;
    C.dt = F.dt;
    C.dx = 2.0f * F.dx;
    C.dof.leaves.clear();
    C.dof.background = -1;
    for (const auto& kv : F.dof.leaves) {
        const Coord fo = kv.first;
        const Coord touched{fo.x / 2, fo.y / 2, fo.z / 2};   // C++ truncating division
        Leaf<int32_t>& nl = C.dof.leaves[leaf_of(touched)];
        (void)nl;
    }
    for (auto& kv : C.dof.leaves) {
        Leaf<int32_t>& lf = kv.second;
        for (int w = 0; w < 8; ++w) lf.on[w] = 0;
        for (int off = 0; off < 512; ++off) {
            const Coord lc = local_of(off);
            const Coord fb{2 * (kv.first.x + lc.x), 2 * (kv.first.y + lc.y), 2 * (kv.first.z + lc.z)};
            bool found = false;
            for (int ii = 0; ii < 2 && !found; ++ii)
                for (int jj = 0; jj < 2 && !found; ++jj)
                    for (int kk = 0; kk < 2 && !found; ++kk)
                        found = F.dof.is_on(Coord{fb.x + ii, fb.y + jj, fb.z + kk});
            lf.set_on(off, found);
        }
    }
    set_dof_index(C);
    const float t = C.dt / (C.dx * C.dx);
    C.diag.topology_copy(C.dof, 6.0f * t);
    C.xe.topology_copy(C.dof, -t);
    C.ye.topology_copy(C.dof, -t);
    C.ze.topology_copy(C.dof, -t);
    for (auto& kv : C.dof.leaves) {
        const Coord o = kv.first;
        Leaf<float>* dl = C.diag.probe(o);
        Leaf<float>* xl = C.xe.probe(o);
        Leaf<float>* yl = C.ye.probe(o);
        Leaf<float>* zl = C.ze.probe(o);
        for (int off = 0; off < 512; ++off) {
            if (!kv.second.is_on(off)) continue;
            const Coord lc = local_of(off);
            const Coord fb{2 * (o.x + lc.x), 2 * (o.y + lc.y), 2 * (o.z + lc.z)};
            float diag = 0, x = 0, y = 0, z = 0;
            for (int ii = 0; ii < 2; ++ii)
                for (int jj = 0; jj < 2; ++jj)
                    for (int kk = 0; kk < 2; ++kk) {
                        const Coord fc{fb.x + ii, fb.y + jj, fb.z + kk};
                        if (!F.dof.is_on(fc)) continue;
                        diag += F.diag.get(fc);
                        if (F.dof.is_on(Coord{fc.x - 1, fc.y, fc.z})) {
                            if (ii == 0) x += F.xe.get(fc);
                            else diag += 2 * F.xe.get(fc);
                        }
                        if (F.dof.is_on(Coord{fc.x, fc.y - 1, fc.z})) {
                            if (jj == 0) y += F.ye.get(fc);
                            else diag += 2 * F.ye.get(fc);
                        }
                        if (F.dof.is_on(Coord{fc.x, fc.y, fc.z - 1})) {
                            if (kk == 0) z += F.ze.get(fc);
                            else diag += 2 * F.ze.get(fc);
                        }
                    }
            const float factor = 0.5f * (1.0f / 8.0f);
            dl->v[off] = diag * factor;
            xl->v[off] = x * factor;
            yl->v[off] = y * factor;
            zl->v[off] = z * factor;
        }
    }
    trim_level(C);
    linearise(C);
}

// init_SPAI0 (:1113-1190)
void build_spai0(Level& L) {
    if (L.spai0_ready) return;
    const float dflt = 6.0f / (7.f * (6.f * L.t));
    L.spai0.topology_copy(L.dof, dflt);
    auto sqr = [](float v) { return v * v; };
    for (auto& kv : L.dof.leaves) {
        Leaf<float>* sl = L.spai0.probe(kv.first);
        for (int off = 0; off < 512; ++off) {
            if (!kv.second.is_on(off)) continue;
            const Coord lc = local_of(off);
            const Coord g{kv.first.x + lc.x, kv.first.y + lc.y, kv.first.z + lc.z};
            float row = 0;
            const FGrid* te[3] = {&L.xe, &L.ye, &L.ze};
            for (int c = 0; c < 3; ++c) {
                Coord up = g, dn = g;
                up[c]++;
                dn[c]--;
                if (L.dof.is_on(up)) row += sqr(te[c]->get(up));
                if (L.dof.is_on(dn)) row += sqr(te[c]->get(g));
            }
            const float d = L.diag.get(g);
            sl->v[off] = d == 0 ? 0.f : d / (sqr(d) + row);
        }
    }
    trim(L.spai0, dflt, dflt * 1e-5f);
    for (size_t i = 0; i < L.nleaf(); ++i) L.cspai[i] = data_or_null(L.spai0, L.org[i]);
    L.spai0_ready = true;
}

// ------------------------------------------------------------ operators
enum Mode { kLap = 0, kRes = 1, kJac = 2, kSpai = 3, kRed = 4, kBlack = 5 };

template <class R>
inline R mul(R a, R b) { return a * b; }

inline float fmaT(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double fmaT(double a, double b, double c) { return std::fma(a, b, c); }

template <class R>
struct Vec {
    std::vector<R> d;   // [nleaf][512] over the linearised DOF-tree leaves
};

// LightWeightApplier::operator() (:755-1031) evaluated voxel by voxel.
template <class R>
void apply(const Level& L, int mode, R w, const R* x, const R* b, R* y) {
    const int64_t n = int64_t(L.nleaf());
    const R one_minus_w = R(1) - w;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const Leaf<int32_t>& m = *L.dl[i];
        const R* self = x + i * 512;
        auto nbr = [&](int f) -> const R* { return L.nb[i][f] >= 0 ? x + L.nb[i][f] * 512 : nullptr; };
        const R *xp = nbr(0), *xm = nbr(1), *yp = nbr(2), *ym = nbr(3), *zp = nbr(4), *zm = nbr(5);
        auto cf = [&](const float* p, int off, float d) -> R { return R(p ? p[off] : d); };
        R out[512];
        for (int row = 0; row < 64; ++row) {
            const int base = row * 8;
            const uint8_t vm = uint8_t(m.on[row >> 3] >> ((row & 7) * 8));
            if (vm == 0) continue;
            const int lx = row >> 3, ly = row & 7;
            for (int z = 0; z < 8; ++z) {
                const int off = base + z;
                // lhs neighbours (absent leaf -> 0)
                const R vxp = lx < 7 ? self[off + 64] : (xp ? xp[off - 448] : R(0));
                const R vxm = lx > 0 ? self[off - 64] : (xm ? xm[off + 448] : R(0));
                const R vyp = ly < 7 ? self[off + 8] : (yp ? yp[off - 56] : R(0));
                const R vym = ly > 0 ? self[off - 8] : (ym ? ym[off + 56] : R(0));
                const R vzp = z < 7 ? self[off + 1] : (zp ? zp[off - 7] : R(0));
                const R vzm = z > 0 ? self[off - 1] : (zm ? zm[off + 7] : R(0));
                // face coefficients (trimmed or absent leaf -> default)
                const R txp = lx < 7 ? cf(L.cx[i], off + 64, L.d_face) : cf(L.cxu[i], off - 448, L.d_face);
                const R txm = cf(L.cx[i], off, L.d_face);
                const R typ = ly < 7 ? cf(L.cy[i], off + 8, L.d_face) : cf(L.cyu[i], off - 56, L.d_face);
                const R tym = cf(L.cy[i], off, L.d_face);
                R tzp;
                if (z < 7) tzp = cf(L.cz[i], off + 1, L.d_face);
                else tzp = cf(L.czu[i], off - 7, L.d_face);
                const R tzm = cf(L.cz[i], off, L.d_face);
                R f1 = mul(vxp, txp), f2 = mul(vxm, txm), f3 = mul(vyp, typ), f4 = mul(vym, tym);
                R f5 = mul(vzp, tzp), f6 = mul(vzm, tzm);
                f1 = f1 + f2;
                f3 = f3 + f4;
                f5 = f5 + f6;
                f1 = f1 + f3;
                f1 = f1 + f5;
                const R lhs = self[off];
                R res;
                switch (mode) {
                    case kJac: {
                        const R ax = fmaT(lhs, cf(L.cdiag[i], off, L.d_diag), f1);
                        R r = b[i * 512 + off] - ax;
                        r = r * cf(L.cinv[i], off, L.d_inv);
                        res = fmaT(w, r, lhs);
                        break;
                    }
                    case kRes: res = b[i * 512 + off] - fmaT(lhs, cf(L.cdiag[i], off, L.d_diag), f1); break;
                    case kSpai: {
                        const R ax = fmaT(lhs, cf(L.cdiag[i], off, L.d_diag), f1);
                        const R r = b[i * 512 + off] - ax;
                        res = fmaT(cf(L.cspai[i], off, L.d_spai), r, lhs);
                        break;
                    }
                    case kRed:
                    case kBlack: {
                        R r = b[i * 512 + off] - f1;
                        r = r * cf(L.cinv[i], off, L.d_inv);
                        r = r * w;
                        const R upd = fmaT(lhs, one_minus_w, r);
                        const bool red = ((lx + ly + z) % 2) == 0;
                        res = (red == (mode == kRed)) ? upd : lhs;
                        break;
                    }
                    default: res = fmaT(lhs, cf(L.cdiag[i], off, L.d_diag), f1); break;
                }
                out[off] = m.is_on(off) ? res : R(0);
            }
        }
        // write after computing the leaf (GS colours only read the other colour)
        for (int row = 0; row < 64; ++row) {
            const uint8_t vm = uint8_t(m.on[row >> 3] >> ((row & 7) * 8));
            if (vm == 0) continue;
            for (int z = 0; z < 8; ++z) y[i * 512 + row * 8 + z] = out[row * 8 + z];
        }
    }
}

// setGridToResultAfterFirstJacobi / SPAI0 / RBGS (:1435-1582); kind 2 runs the
// red update from zero followed by the black pass.
template <class R>
void first_sweep(const Level& L, int kind, R w, const R* b, R* y) {
    const int64_t n = int64_t(L.nleaf());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const Leaf<int32_t>& m = *L.dl[i];
        for (int row = 0; row < 64; ++row) {
            const uint8_t vm = uint8_t(m.on[row >> 3] >> ((row & 7) * 8));
            const bool skip_row = vm == 0 && kind == 0;   // the Jacobi variant skips empty rows
            if (skip_row) continue;
            for (int z = 0; z < 8; ++z) {
                const int off = row * 8 + z;
                const R bv = b[i * 512 + off];
                R v;
                if (kind == 1) {
                    v = R(L.cspai[i] ? L.cspai[i][off] : L.d_spai) * bv;
                } else {
                    v = R(L.cinv[i] ? L.cinv[i][off] : L.d_inv) * bv;
                    v = w * v;
                    if (kind == 2 && (((row >> 3) + (row & 7) + z) % 2) != 0) v = R(0);
                }
                y[i * 512 + off] = (kind == 0 || m.is_on(off)) ? v : R(0);
            }
        }
    }
    if (kind == 2) apply<R>(L, kBlack, w, y, b, y);
}

// restriction (:1624-1684): coarse = 1/8 * sum of on fine children
template <class R>
void restrict_to(const Level& F, const Level& C, const R* fine, R* coarse) {
    std::map<Coord, int64_t, TreeLess> fpos;
    for (size_t i = 0; i < F.nleaf(); ++i) fpos[F.org[i]] = int64_t(i);
    const int64_t n = int64_t(C.nleaf());
#pragma omp parallel for schedule(static)
    for (int64_t ci = 0; ci < n; ++ci) {
        const Coord co = C.org[ci];
        int64_t fl[8];
        int k = 0;
        for (int ii = 0; ii < 16; ii += 8)
            for (int jj = 0; jj < 16; jj += 8)
                for (int kk = 0; kk < 16; kk += 8) {
                    auto it = fpos.find(Coord{2 * co.x + ii, 2 * co.y + jj, 2 * co.z + kk});
                    fl[k++] = it == fpos.end() ? -1 : it->second;
                }
        const Leaf<int32_t>& cm = *C.dl[ci];
        for (int off = 0; off < 512; ++off) {
            if (!cm.is_on(off)) continue;
            const Coord lc = local_of(off);
            const int which = (lc.z >= 4 ? 1 : 0) + (lc.y >= 4 ? 2 : 0) + (lc.x >= 4 ? 4 : 0);
            if (fl[which] < 0) continue;
            const Leaf<int32_t>& fm = *F.dl[fl[which]];
            const Coord gfine{2 * (co.x + lc.x), 2 * (co.y + lc.y), 2 * (co.z + lc.z)};
            const int fb = offset_of(gfine);
            R sum = 0;
            for (int ii = 0; ii < 2; ++ii)
                for (int jj = 0; jj < 2; ++jj)
                    for (int kk = 0; kk < 2; ++kk) {
                        const int fo = fb + 64 * ii + 8 * jj + kk;
                        if (fm.is_on(fo)) sum += fine[fl[which] * 512 + fo];
                    }
            coarse[ci * 512 + off] = sum * R(0.125f);
        }
    }
}

// prolongation<true> (:1686-1754): fine += alpha * coarse on on fine voxels
template <class R>
void prolong_add(const Level& F, const Level& C, const R* coarse, R* fine, R alpha) {
    std::map<Coord, int64_t, TreeLess> fpos;
    for (size_t i = 0; i < F.nleaf(); ++i) fpos[F.org[i]] = int64_t(i);
    const int64_t n = int64_t(C.nleaf());
#pragma omp parallel for schedule(static)
    for (int64_t ci = 0; ci < n; ++ci) {
        const Coord co = C.org[ci];
        int64_t fl[8];
        int k = 0;
        for (int ii = 0; ii < 16; ii += 8)
            for (int jj = 0; jj < 16; jj += 8)
                for (int kk = 0; kk < 16; kk += 8) {
                    auto it = fpos.find(Coord{2 * co.x + ii, 2 * co.y + jj, 2 * co.z + kk});
                    fl[k++] = it == fpos.end() ? -1 : it->second;
                }
        const Leaf<int32_t>& cm = *C.dl[ci];
        for (int off = 0; off < 512; ++off) {
            if (!cm.is_on(off)) continue;
            const Coord lc = local_of(off);
            const int which = (lc.z >= 4 ? 1 : 0) + (lc.y >= 4 ? 2 : 0) + (lc.x >= 4 ? 4 : 0);
            if (fl[which] < 0) continue;
            const Leaf<int32_t>& fm = *F.dl[fl[which]];
            const R cv = alpha * coarse[ci * 512 + off];
            const int fb = offset_of(Coord{2 * (co.x + lc.x), 2 * (co.y + lc.y), 2 * (co.z + lc.z)});
            for (int ii = 0; ii < 2; ++ii)
                for (int jj = 0; jj < 2; ++jj)
                    for (int kk = 0; kk < 2; ++kk) {
                        const int fo = fb + (ii ? 64 : 0) + (jj ? 8 : 0) + kk;
                        if (fm.is_on(fo)) fine[fl[which] * 512 + fo] += cv;
                    }
        }
    }
}

// dot over on voxels: per-leaf partials in R, leaves summed in order (grid_dot_op)
template <class R>
R dot(const Level& L, const R* a, const R* b) {
    const int64_t n = int64_t(L.nleaf());
    std::vector<R> part(n, R(0));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        R s = 0;
        const Leaf<int32_t>& m = *L.dl[i];
        for (int off = 0; off < 512; ++off)
            if (m.is_on(off)) s += a[i * 512 + off] * b[i * 512 + off];
        part[i] = s;
    }
    R t = 0;
    for (int64_t i = 0; i < n; ++i) t += part[i];
    return t;
}

// grid_abs_max_op (openvdb_grid_math_op.h:4-40): the first non-finite value wins
template <class R>
R absmax(const Level& L, const R* a) {
    R m = 0;
    for (size_t i = 0; i < L.nleaf(); ++i) {
        const Leaf<int32_t>& lf = *L.dl[i];
        for (int off = 0; off < 512; ++off) {
            if (!lf.is_on(off)) continue;
            const R v = a[i * 512 + off];
            if (!std::isfinite(v)) return v;
            m = std::max(m, std::abs(v));
        }
    }
    return m;
}

template <class R>
void axpy(const Level& L, R alpha, const R* x, R* y) {   // y += alpha x (levelAlphaXPlusY)
    const int64_t n = int64_t(L.nleaf());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int off = 0; off < 512; ++off)
            if (L.dl[i]->is_on(off)) {
                const R t = alpha * x[i * 512 + off];
                y[i * 512 + off] = y[i * 512 + off] + t;
            }
}

template <class R>
void xpay(const Level& L, R alpha, const R* x, R* y) {   // y = x + alpha y (levelXPlusAlphaY)
    const int64_t n = int64_t(L.nleaf());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int off = 0; off < 512; ++off)
            if (L.dl[i]->is_on(off)) {
                const R t = alpha * y[i * 512 + off];
                y[i * 512 + off] = x[i * 512 + off] + t;
            }
}

// ------------------------------------------------------- coarsest solve
// Explicit rows of the coarsest level (getTriplets :129-206) and
// Eigen::ConjugateGradient<SparseMatrix<float>> with the diagonal
// preconditioner, zero guess, 10 iterations, tolerance = epsilon
// (constructCoarsestLevelExactSolver :2114-2126). Eigen's own summation order
// inside its SpMV and dot products is parity unpinned; the algorithm and the
// stopping rule follow ConjugateGradient.h (conjugate_gradient()).
struct Coarsest {
    int64_t n = 0;
    std::vector<int64_t> pos;                  // linear brick index
    std::vector<std::array<int64_t, 6>> nb;    // +x,-x,+y,-y,+z,-z DOF or -1
    std::vector<std::array<float, 8>> c;       // diag, c+x, c-x, c+y, c-y, c+z, c-z, invdiag
};

void build_coarsest(const Level& L, Coarsest& K) {
    K.n = L.ndof;
    K.pos.assign(K.n, 0);
    K.nb.assign(K.n, {-1, -1, -1, -1, -1, -1});
    K.c.assign(K.n, {0, 0, 0, 0, 0, 0, 0, 0});
    const FGrid* te[3] = {&L.xe, &L.ye, &L.ze};
    for (size_t i = 0; i < L.nleaf(); ++i) {
        const Leaf<int32_t>& lf = *L.dl[i];
        for (int off = 0; off < 512; ++off) {
            if (!lf.is_on(off)) continue;
            const int64_t d = lf.v[off];
            const Coord lc = local_of(off);
            const Coord g{L.org[i].x + lc.x, L.org[i].y + lc.y, L.org[i].z + lc.z};
            K.pos[d] = int64_t(i) * 512 + off;
            const float dg = L.diag.get(g);
            K.c[d][0] = dg;
            for (int c = 0; c < 3; ++c) {
                Coord up = g, dn = g;
                up[c]++;
                dn[c]--;
                if (L.dof.is_on(up)) {
                    K.nb[d][2 * c] = L.dof.get(up);
                    K.c[d][1 + 2 * c] = te[c]->get(up);
                }
                if (L.dof.is_on(dn)) {
                    K.nb[d][2 * c + 1] = L.dof.get(dn);
                    K.c[d][2 + 2 * c] = te[c]->get(g);
                }
            }
            K.c[d][7] = dg != 0.f ? 1.0f / dg : 1.0f;
        }
    }
}

template <class R>
void coarsest_solve(const Coarsest& K, const R* bricks_b, R* bricks_x) {
    const int64_t n = K.n;
    std::vector<R> x(n, R(0)), r(n), p(n), z(n), q(n);
    for (int64_t d = 0; d < n; ++d) r[d] = bricks_b[K.pos[d]];
    auto dotv = [&](const std::vector<R>& a, const std::vector<R>& b) {
        R s = 0;
        for (int64_t d = 0; d < n; ++d) s += a[d] * b[d];
        return s;
    };
    auto spmv = [&](const std::vector<R>& v, std::vector<R>& out) {
        for (int64_t d = 0; d < n; ++d) {
            R nbv[6];
            for (int k = 0; k < 6; ++k) nbv[k] = K.nb[d][k] >= 0 ? v[K.nb[d][k]] : R(0);
            R f1 = nbv[0] * R(K.c[d][1]), f2 = nbv[1] * R(K.c[d][2]), f3 = nbv[2] * R(K.c[d][3]);
            R f4 = nbv[3] * R(K.c[d][4]), f5 = nbv[4] * R(K.c[d][5]), f6 = nbv[5] * R(K.c[d][6]);
            out[d] = fmaT(v[d], R(K.c[d][0]), ((f1 + f2) + (f3 + f4)) + (f5 + f6));
        }
    };
    const R tol = std::is_same<R, float>::value ? R(FLT_EPSILON) : R(DBL_EPSILON);
    const R rhs_norm2 = dotv(r, r);
    bool done = rhs_norm2 == R(0);
    const R threshold = std::max(R(tol * tol * rhs_norm2), std::is_same<R, float>::value ? R(FLT_MIN) : R(DBL_MIN));
    if (!done && rhs_norm2 < threshold) done = true;
    if (!done) {
        for (int64_t d = 0; d < n; ++d) p[d] = R(K.c[d][7]) * r[d];
        R abs_new = dotv(r, p);
        for (int it = 0; it < 10; ++it) {
            spmv(p, q);
            const R alpha = abs_new / dotv(p, q);
            for (int64_t d = 0; d < n; ++d) {
                x[d] = x[d] + alpha * p[d];
                r[d] = r[d] - alpha * q[d];
            }
            if (dotv(r, r) < threshold) break;
            for (int64_t d = 0; d < n; ++d) z[d] = R(K.c[d][7]) * r[d];
            const R abs_old = abs_new;
            abs_new = dotv(r, z);
            const R beta = abs_new / abs_old;
            for (int64_t d = 0; d < n; ++d) p[d] = z[d] + beta * p[d];
        }
    }
    for (int64_t d = 0; d < n; ++d) bricks_x[K.pos[d]] = x[d];
}

// -------------------------------------------------------------- solver
enum Smoother { kSched = 0, kWJ = 1, kRBGS = 2, kSPAI = 3 };

float sched_pc(int n, int i) {   // muCyclePreconditioner::get_scheduled_weight
    if (i >= n) return 6.0f / 7.0f;
    if (n == 1) return 6.0f / 7.0f;
    if (n == 2) return i == 0 ? 1.7319f : 0.5695f;
    const float w[3] = {2.2473f, 0.8571f, 0.5296f};
    return w[i];
}
float sched_it(int n, int i) {   // muCycleIterative::get_scheduled_weight
    if (i >= n) return 6.0f / 7.0f;
    if (n == 2) return i == 0 ? 1.7319f : 0.5695f;
    if (n >= 3 && i < 3) {
        const float w[3] = {2.2473f, 0.8571f, 0.5296f};
        return w[i];
    }
    return 6.0f / 7.0f;
}

template <class R>
struct Solver {
    std::vector<Level>* levels = nullptr;
    const Coarsest* K = nullptr;
    int smoother = kSched;
    std::vector<float> sor;
    // per-level scratch: lhs, rhs, temp (mMuCycleLHSs / RHSs / Temps)
    std::vector<std::vector<R>> lhs, rhs, tmp;

    void init(std::vector<Level>& lv, const Coarsest& k) {
        levels = &lv;
        K = &k;
        const size_t L = lv.size();
        sor.assign(L, 1.2f);
        lhs.resize(L);
        rhs.resize(L);
        tmp.resize(L);
        for (size_t l = 0; l < L; ++l) {
            lhs[l].assign(lv[l].nleaf() * 512, R(0));
            rhs[l].assign(lv[l].nleaf() * 512, R(0));
            tmp[l].assign(lv[l].nleaf() * 512, R(0));
        }
    }
    Level& lv(int l) { return (*levels)[l]; }
    int nl() const { return int(levels->size()); }

    // one smoothing sweep on (cur, oth): Jacobi-type sweeps write oth and swap
    void sweep(int l, R*& cur, R*& oth, const R* b, float w_jacobi, bool red_first, bool iterative) {
        Level& L = lv(l);
        switch (smoother) {
            default:
            case kSched:
            case kWJ:
                apply<R>(L, kJac, R(w_jacobi), cur, b, oth);
                std::swap(cur, oth);
                break;
            case kRBGS:
                if (iterative) sor[l] = 1.0f;
                apply<R>(L, red_first ? kRed : kBlack, R(sor[l]), cur, b, cur);
                apply<R>(L, red_first ? kBlack : kRed, R(sor[l]), cur, b, cur);
                break;
            case kSPAI:
                apply<R>(L, kSpai, R(0), cur, b, oth);
                std::swap(cur, oth);
                break;
        }
    }

    // muCyclePreconditioner<2, skip_first> with n = 3 (:1844-1977)
    void precond(int l, R* X, const R* B, bool skip_first) {
        const int n = 3;
        if (l == nl() - 1) {
            coarsest_solve<R>(*K, B, X);
            return;
        }
        Level& L = lv(l);
        R* cur = X;
        R* oth = tmp[l].data();
        if (skip_first) {
            switch (smoother) {
                default:
                case kSched:
                    first_sweep<R>(L, 0, R(sched_pc(n, 0)), B, oth);
                    std::swap(cur, oth);
                    break;
                case kWJ:
                    first_sweep<R>(L, 0, R(6.0f / 7.0f), B, oth);
                    std::swap(cur, oth);
                    break;
                case kRBGS: first_sweep<R>(L, 2, R(sor[l]), B, cur); break;
                case kSPAI:
                    first_sweep<R>(L, 1, R(0), B, oth);
                    std::swap(cur, oth);
                    break;
            }
        }
        for (int i = skip_first ? 1 : 0; i < n; ++i)
            sweep(l, cur, oth, B, smoother == kWJ ? 6.0f / 7.0f : sched_pc(n, i), true, false);
        apply<R>(L, kRes, R(0), cur, B, oth);
        restrict_to<R>(L, lv(l + 1), oth, rhs[l + 1].data());
        precond(l + 1, lhs[l + 1].data(), rhs[l + 1].data(), true);
        precond(l + 1, lhs[l + 1].data(), rhs[l + 1].data(), false);
        prolong_add<R>(L, lv(l + 1), lhs[l + 1].data(), cur, R(1));
        for (int i = 0; i < n; ++i)
            sweep(l, cur, oth, B, smoother == kWJ ? 6.0f / 7.0f : sched_pc(n, n - 1 - i), false, false);
    }

    // muCycleIterative<2> with n = 3 (:1979-2110)
    void mg_cycle(int l, R* X, const R* B) {
        const int n = 3;
        R* cur = X;
        R* oth = tmp[l].data();
        if (l == nl() - 1) {
            for (int i = 0; i < 10 * n; ++i)
                sweep(l, cur, oth, B, smoother == kWJ ? 6.0f / 7.0f : sched_it(n, i), true, true);
            return;
        }
        Level& L = lv(l);
        for (int i = 0; i < n; ++i) sweep(l, cur, oth, B, smoother == kWJ ? 6.0f / 7.0f : sched_it(n, i), true, true);
        apply<R>(L, kRes, R(0), cur, B, oth);
        restrict_to<R>(L, lv(l + 1), oth, rhs[l + 1].data());
        for (int mu = 0; mu < 2; ++mu) mg_cycle(l + 1, lhs[l + 1].data(), rhs[l + 1].data());
        prolong_add<R>(L, lv(l + 1), lhs[l + 1].data(), cur, R(0.5f));
        for (int i = 0; i < n; ++i)
            sweep(l, cur, oth, B, smoother == kWJ ? 6.0f / 7.0f : sched_it(n, n - 1 - i), false, true);
    }

    // solveMultigridPCG (:2155-2226)
    int solve_pcg(R* x, const R* b, R tol, int maxit, int* iters, std::vector<R>* history) {
        Level& L0 = lv(0);
        const size_t N = L0.nleaf() * 512;
        std::vector<R> r(N, R(0)), p(N, R(0)), z(N, R(0));
        int it = 0;
        apply<R>(L0, kRes, R(0), x, b, r.data());
        R nu = absmax<R>(L0, r.data());
        const R numax = tol * nu;
        if (history) history->push_back(nu);
        if (nu <= numax) {
            *iters = 0;
            return 0;
        }
        precond(0, p.data(), r.data(), true);
        R rho = dot<R>(L0, p.data(), r.data());
        R nu_old = nu;
        for (; it < maxit; ++it) {
            apply<R>(L0, kLap, R(0), p.data(), nullptr, z.data());
            const R sigma = dot<R>(L0, p.data(), z.data());
            const R alpha = rho / sigma;
            axpy<R>(L0, -alpha, z.data(), r.data());
            nu_old = nu;
            nu = absmax<R>(L0, r.data());
            if (history) history->push_back(nu);
            if (nu <= numax) {
                axpy<R>(L0, alpha, p.data(), x);
                *iters = it;
                return 0;
            }
            if (nu > nu_old && it > 3) {
                *iters = it;
                return 1;
            }
            precond(0, z.data(), r.data(), true);
            const R rho_new = dot<R>(L0, z.data(), r.data());
            const R beta = rho_new / rho;
            rho = rho_new;
            axpy<R>(L0, alpha, p.data(), x);
            xpay<R>(L0, beta, z.data(), p.data());
        }
        *iters = it;
        return 1;
    }

    // solvePureMultigrid (:2228-2267)
    int solve_mg(R* x, const R* b, R tol, int maxit, int* iters) {
        Level& L0 = lv(0);
        std::vector<R> r(L0.nleaf() * 512, R(0));
        apply<R>(L0, kRes, R(0), x, b, r.data());
        R nu = absmax<R>(L0, r.data());
        const R numax = tol * nu;
        int it = 0;
        if (nu <= numax) {
            *iters = 0;
            return 0;
        }
        for (; it < maxit; ++it) {
            mg_cycle(0, x, b);
            apply<R>(L0, kRes, R(0), x, b, r.data());
            const R nu_old = nu;
            nu = absmax<R>(L0, r.data());
            if (nu <= numax) {
                *iters = it;
                return 0;
            }
            if (nu > nu_old && it > 8) {
                *iters = it;
                return 1;
            }
        }
        *iters = it;
        return 1;
    }
};

struct Handle {
    std::vector<Level> levels;
    Coarsest K;
    Solver<float> s32;
    Solver<double> s64;
    bool s64_ready = false;
};

template <class T>
void load_grid(int64_t n, const int32_t* org, const uint64_t* act, const float* val, int ncomp, const float* bg,
               Grid<T>& g);

template <>
void load_grid<float>(int64_t n, const int32_t* org, const uint64_t* act, const float* val, int, const float* bg,
                      FGrid& g) {
    g.background = bg ? bg[0] : 0.f;
    for (int64_t i = 0; i < n; ++i) {
        Leaf<float>& l = g.leaves[Coord{org[3 * i], org[3 * i + 1], org[3 * i + 2]}];
        for (int w = 0; w < 8; ++w) l.on[w] = act ? act[8 * i + w] : ~uint64_t(0);
        std::memcpy(l.v, val + 512 * i, 512 * sizeof(float));
    }
}

template <>
void load_grid<Vec3>(int64_t n, const int32_t* org, const uint64_t* act, const float* val, int, const float* bg,
                     VGrid& g) {
    for (int c = 0; c < 3; ++c) g.background.v[c] = bg ? bg[c] : 0.f;
    for (int64_t i = 0; i < n; ++i) {
        Leaf<Vec3>& l = g.leaves[Coord{org[3 * i], org[3 * i + 1], org[3 * i + 2]}];
        for (int w = 0; w < 8; ++w) l.on[w] = act ? act[8 * i + w] : ~uint64_t(0);
        for (int k = 0; k < 512; ++k)
            for (int c = 0; c < 3; ++c) l.v[k].v[c] = val[(512 * i + k) * 3 + c];
    }
}

// DOF-ordered vector <-> bricks of a level
template <class R, class S>
void dof_to_bricks(const Level& L, const S* dv, std::vector<R>& br) {
    br.assign(L.nleaf() * 512, R(0));
    for (size_t i = 0; i < L.nleaf(); ++i)
        for (int off = 0; off < 512; ++off)
            if (L.dl[i]->is_on(off)) br[i * 512 + off] = R(dv[L.dl[i]->v[off]]);
}
template <class R, class S>
void bricks_to_dof(const Level& L, const std::vector<R>& br, S* dv) {
    for (size_t i = 0; i < L.nleaf(); ++i)
        for (int off = 0; off < 512; ++off)
            if (L.dl[i]->is_on(off)) dv[L.dl[i]->v[off]] = S(br[i * 512 + off]);
}

}  // namespace orc

using namespace orc;

extern "C" {

void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}
int orc_get_threads(void) { return omp_get_max_threads(); }

float orc_fraction_inside2(float a, float b) { return fraction_inside(a, b); }
float orc_fraction_inside4(float bl, float br, float tl, float tr) { return fraction_inside4(bl, br, tl, tr); }

void* orc_poisson_create(int64_t nphi, const int32_t* phi_org, const uint64_t* phi_act, const float* phi_val,
                         float phi_bg, int64_t nw, const int32_t* w_org, const uint64_t* w_act, const float* w_val,
                         const float* w_bg, float dt, float dx, int64_t max_coarsest) {
    auto* h = new Handle();
    FGrid phi;
    VGrid w;
    load_grid<float>(nphi, phi_org, phi_act, phi_val, 1, &phi_bg, phi);
    load_grid<Vec3>(nw, w_org, w_act, w_val, 3, w_bg, w);
    h->levels.reserve(32);
    h->levels.emplace_back();
    build_finest(h->levels.back(), phi, w, dt, dx);
    while (h->levels.back().ndof > max_coarsest) {
        h->levels.emplace_back();
        build_coarse(h->levels.back(), h->levels[h->levels.size() - 2]);
    }
    build_coarsest(h->levels.back(), h->K);
    h->s32.init(h->levels, h->K);
    return h;
}

void orc_poisson_destroy(void* p) { delete static_cast<Handle*>(p); }

int orc_nlevels(void* p) { return int(static_cast<Handle*>(p)->levels.size()); }

int64_t orc_level_ndof(void* p, int l) { return static_cast<Handle*>(p)->levels[l].ndof; }
int64_t orc_level_nleaf_tree(void* p, int l) { return int64_t(static_cast<Handle*>(p)->levels[l].nleaf()); }

// leaves holding at least one DOF, in tree order
int64_t orc_level_nleaf_dof(void* p, int l) {
    const Level& L = static_cast<Handle*>(p)->levels[l];
    int64_t n = 0;
    for (size_t i = 0; i < L.nleaf(); ++i) n += L.dl[i]->count() > 0;
    return n;
}

void orc_level_topology(void* p, int l, int32_t* origins, uint64_t* masks, int64_t* dofbase) {
    const Level& L = static_cast<Handle*>(p)->levels[l];
    int64_t k = 0;
    for (size_t i = 0; i < L.nleaf(); ++i) {
        const Leaf<int32_t>& lf = *L.dl[i];
        if (lf.count() == 0) continue;
        if (origins) {
            origins[3 * k] = L.org[i].x;
            origins[3 * k + 1] = L.org[i].y;
            origins[3 * k + 2] = L.org[i].z;
        }
        if (masks) std::memcpy(masks + 8 * k, lf.on, sizeof(lf.on));
        if (dofbase) {
            int32_t first = -1;
            for (int off = 0; off < 512 && first < 0; ++off)
                if (lf.is_on(off)) first = lf.v[off];
            dofbase[k] = first;
        }
        ++k;
    }
}

// effective coefficient bricks over the leaves holding DOFs, and whether each leaf survived trimming
void orc_level_coefficients(void* p, int l, int kind, float* bricks, uint8_t* kept) {
    Handle* h = static_cast<Handle*>(p);
    Level& L = h->levels[l];
    if (kind == 4) build_spai0(L);
    const FGrid* g = kind == 0 ? &L.diag : kind == 1 ? &L.xe : kind == 2 ? &L.ye : kind == 3 ? &L.ze : &L.spai0;
    const float dflt = kind == 0 ? L.d_diag : kind == 4 ? L.d_spai : L.d_face;
    int64_t k = 0;
    for (size_t i = 0; i < L.nleaf(); ++i) {
        const Leaf<int32_t>& lf = *L.dl[i];
        if (lf.count() == 0) continue;
        const Leaf<float>* cl = g->probe(L.org[i]);
        if (kept) kept[k] = cl != nullptr;
        if (bricks)
            for (int off = 0; off < 512; ++off) bricks[k * 512 + off] = cl ? cl->v[off] : dflt;
        ++k;
    }
}

// BuildPoissonRhs (simd_vdb_poisson.cpp:19-93, driven by
// createPressurePoissonRightHandSide :95-127): per DOF voxel, over the faces
// x+ x- y+ y- z+ z- (face f of channel c read at the voxel, +1 along c for the
// + face), rhs -= / += invdx * (w * u + (1 - w) * u_solid); a voxel whose six
// face weights are all 0 gets 0. The solid velocity is a Vec3 grid like the
// weights (ns = 0: no leaves, every lookup reads s_bg). Pinned by
// tests/test_oracle_rhs.py (closed forms of uniform, linear, quadratic and
// blended fields, and a dense brute-force divergence).
void orc_build_rhs(void* p, int64_t nw, const int32_t* w_org, const float* w_val, const float* w_bg, int64_t nv,
                   const int32_t* v_org, const float* vx, const float* vy, const float* vz, float v_bg, int64_t ns,
                   const int32_t* s_org, const float* s_val, const float* s_bg, float* rhs) {
    Handle* h = static_cast<Handle*>(p);
    const Level& L = h->levels[0];
    VGrid w, sv;
    FGrid v[3];
    load_grid<Vec3>(nw, w_org, nullptr, w_val, 3, w_bg, w);
    load_grid<Vec3>(ns, s_org, nullptr, s_val, 3, s_bg, sv);
    const float* vv[3] = {vx, vy, vz};
    for (int c = 0; c < 3; ++c) load_grid<float>(nv, v_org, nullptr, vv[c], 1, &v_bg, v[c]);
    const float invdx = 1.0f / L.dx;
    for (size_t i = 0; i < L.nleaf(); ++i) {
        const Leaf<int32_t>& lf = *L.dl[i];
        for (int off = 0; off < 512; ++off) {
            if (!lf.is_on(off)) continue;
            const Coord lc = local_of(off);
            const Coord g{L.org[i].x + lc.x, L.org[i].y + lc.y, L.org[i].z + lc.z};
            float r = 0;
            bool nz = false;
            for (int f = 0; f < 6; ++f) {
                const int c = f / 2;
                const bool pos = f % 2 == 0;
                Coord nc = g;
                nc[c] += int(pos);
                const float wt = w.get(nc).v[c];
                if (wt != 0.f) nz = true;
                const float vel = v[c].get(nc);
                const float svel = sv.get(nc).v[c];
                const float term = invdx * (wt * vel + (1.0f - wt) * svel);
                if (pos) r -= term;
                else r += term;
            }
            rhs[lf.v[off]] = nz ? r : 0.f;
        }
    }
}

// FLIP_Solver_OpenVDB::calculate_face_weights (FLIP_vdb.cpp:1990-2068): the
// liquid topology dilated twice over the 26 neighbours (voxel by voxel: every
// voxel within Chebyshev distance 2 of an active one), weights from the solid
// SDF at the face corners; inactive voxels hold 1. cut > 0 restates the member
// variant FLIP's substep uses (FLIP_vdb.cpp:2669-2752): after the clamp, a
// weight below cut (1e-2 there) becomes 0. Two-phase like the library: *n
// receives the leaf count, the arrays (tree order) are filled when non-null.
void orc_face_weights(int64_t nl, const int32_t* l_org, const uint64_t* l_act, int64_t ns, const int32_t* s_org,
                      const float* s_val, float s_bg, float cut, int64_t* n, int32_t* org, uint64_t* act,
                      float* val) {
    std::map<Coord, Leaf<Vec3>, TreeLess> out;
    Vec3 one;
    one.v[0] = one.v[1] = one.v[2] = 1.f;
    auto touch_leaf = [&](Coord o) -> Leaf<Vec3>& {
        auto it = out.find(o);
        if (it != out.end()) return it->second;
        Leaf<Vec3>& l = out[o];
        for (int i = 0; i < 512; ++i) l.v[i] = one;
        return l;
    };
    for (int64_t i = 0; i < nl; ++i) touch_leaf(Coord{l_org[3 * i], l_org[3 * i + 1], l_org[3 * i + 2]});
    for (int64_t i = 0; i < nl; ++i) {
        const Coord o{l_org[3 * i], l_org[3 * i + 1], l_org[3 * i + 2]};
        for (int off = 0; off < 512; ++off) {
            if (!((l_act[8 * i + (off >> 6)] >> (off & 63)) & 1u)) continue;
            const Coord lc = local_of(off);
            for (int dx = -2; dx <= 2; ++dx)
                for (int dy = -2; dy <= 2; ++dy)
                    for (int dz = -2; dz <= 2; ++dz) {
                        const Coord g{o.x + lc.x + dx, o.y + lc.y + dy, o.z + lc.z + dz};
                        touch_leaf(leaf_of(g)).set_on(offset_of(g), true);
                    }
        }
    }
    *n = int64_t(out.size());
    if (!org || !act || !val) return;
    FGrid solid;
    load_grid<float>(ns, s_org, nullptr, s_val, 1, &s_bg, solid);
    size_t k = 0;
    for (auto& kv : out) {
        Leaf<Vec3>& l = kv.second;
        for (int off = 0; off < 512; ++off) {
            if (!l.is_on(off)) continue;
            const Coord lc = local_of(off);
            const Coord g{kv.first.x + lc.x, kv.first.y + lc.y, kv.first.z + lc.z};
            float sd[2][2][2];
            for (int ii = 0; ii < 2; ++ii)
                for (int jj = 0; jj < 2; ++jj)
                    for (int kk = 0; kk < 2; ++kk) sd[ii][jj][kk] = solid.get(Coord{g.x + ii, g.y + jj, g.z + kk});
            float w[3];
            w[0] = 1.0f - fraction_inside4(sd[0][0][0], sd[0][1][0], sd[0][0][1], sd[0][1][1]);
            w[1] = 1.0f - fraction_inside4(sd[0][0][0], sd[0][0][1], sd[1][0][0], sd[1][0][1]);
            w[2] = 1.0f - fraction_inside4(sd[0][0][0], sd[1][0][0], sd[0][1][0], sd[1][1][0]);
            for (int c = 0; c < 3; ++c) {
                l.v[off].v[c] = std::max(0.f, std::min(w[c], 1.f));
                if (l.v[off].v[c] < cut) l.v[off].v[c] = 0.f;
            }
        }
        org[3 * k] = kv.first.x;
        org[3 * k + 1] = kv.first.y;
        org[3 * k + 2] = kv.first.z;
        std::memcpy(act + 8 * k, l.on, sizeof(l.on));
        for (int off = 0; off < 512; ++off)
            for (int c = 0; c < 3; ++c) val[(k * 512 + size_t(off)) * 3 + size_t(c)] = l.v[off].v[c];
        ++k;
    }
}

// FLIP_Solver_OpenVDB::apply_pressure_gradient (FLIP_vdb.cpp:2571-2667), one
// channel c: the pressure grid is the level-0 DOF topology (TopologyCopy of
// the DOF index grid, 0 off the DOFs) holding the DOF-ordered pressure. The
// velocity leaves come in the caller's order; act / val are updated in place,
// upd / refl receive the update and reflected grids (the incoming velocity
// where no update happens, as after the reference's deepCopy).
void orc_apply_pressure_gradient(void* p, const float* pressure, int64_t nphi, const int32_t* phi_org,
                                 const float* phi_val, float phi_bg, int64_t nw, const int32_t* w_org,
                                 const float* w_val, const float* w_bg, int64_t ns, const int32_t* s_org,
                                 const float* s_val, const float* s_bg, int c, int64_t nv, const int32_t* v_org,
                                 uint64_t* act, float* val, float* upd, float* refl, float dt, float friction) {
    Handle* h = static_cast<Handle*>(p);
    const Level& L = h->levels[0];
    FGrid phi;
    VGrid w, sv;
    load_grid<float>(nphi, phi_org, nullptr, phi_val, 1, &phi_bg, phi);
    load_grid<Vec3>(nw, w_org, nullptr, w_val, 3, w_bg, w);
    const float zero3[3] = {0.f, 0.f, 0.f};
    load_grid<Vec3>(ns, s_org, nullptr, s_val, 3, s_bg ? s_bg : zero3, sv);
    auto p_on = [&](Coord g) { return L.dof.is_on(g); };
    auto p_val = [&](Coord g) { return L.dof.is_on(g) ? pressure[L.dof.get(g)] : 0.f; };
    for (int64_t i = 0; i < nv; ++i) {
        const Coord o{v_org[3 * i], v_org[3 * i + 1], v_org[3 * i + 2]};
        for (int off = 0; off < 512; ++off) {
            float& u = val[i * 512 + off];
            const float u0 = u;
            upd[i * 512 + off] = u0;
            refl[i * 512 + off] = u0;
            uint64_t& word = act[i * 8 + (off >> 6)];
            const uint64_t bit = uint64_t(1) << (off & 63);
            if (!(word & bit)) continue;
            const Coord lc = local_of(off);
            const Coord g{o.x + lc.x, o.y + lc.y, o.z + lc.z};
            bool has_update = false;
            float updated = u0, reflected = u0, du = 0.f;
            const float weight = w.get(g).v[c];
            Coord lower = g;
            lower[c] -= 1;
            if (weight > 0.0f) {
                if (p_on(g) || p_on(lower)) has_update = true;
                if (has_update) {
                    const float phi_this = phi.get(g), phi_below = phi.get(lower);
                    const float p_this = p_val(g), p_below = p_val(lower);
                    float theta = 1.0f;
                    if (phi_this >= 0 || phi_below >= 0) {
                        theta = fraction_inside(phi_below, phi_this);
                        if (theta < 0.02f) theta = 0.02f;
                    }
                    du = -dt * (p_this - p_below) / L.dx / theta;
                    updated += du;
                    reflected = updated + du;
                    if (weight < 1.0f) {
                        const float svel = sv.get(g).v[c];
                        const float sf = (1.0f - weight) * friction;
                        updated = (1.0f - sf) * updated + sf * svel;
                    }
                }
            }
            if (has_update) {
                u = updated;
                upd[i * 512 + off] = du;
                refl[i * 512 + off] = reflected;
            } else {
                word &= ~bit;
            }
        }
    }
}

int orc_apply(void* p, int l, int mode, float w, const float* x, const float* b, float* y) {
    Handle* h = static_cast<Handle*>(p);
    Level& L = h->levels[l];
    if (mode == kSpai) build_spai0(L);
    std::vector<float> xb, bb, yb;
    dof_to_bricks<float>(L, x, xb);
    if (b) dof_to_bricks<float>(L, b, bb);
    else bb.assign(xb.size(), 0.f);
    yb.assign(xb.size(), 0.f);
    if (mode == kRed || mode == kBlack) {
        apply<float>(L, mode, w, xb.data(), bb.data(), xb.data());
        bricks_to_dof<float>(L, xb, y);
    } else {
        apply<float>(L, mode, w, xb.data(), bb.data(), yb.data());
        bricks_to_dof<float>(L, yb, y);
    }
    return 0;
}

int orc_first_sweep(void* p, int l, int kind, float w, const float* b, float* y) {
    Handle* h = static_cast<Handle*>(p);
    Level& L = h->levels[l];
    if (kind == 1) build_spai0(L);
    std::vector<float> bb, yb;
    dof_to_bricks<float>(L, b, bb);
    yb.assign(bb.size(), 0.f);
    first_sweep<float>(L, kind, w, bb.data(), yb.data());
    bricks_to_dof<float>(L, yb, y);
    return 0;
}

int orc_restrict(void* p, int fl, const float* fine, float* coarse) {
    Handle* h = static_cast<Handle*>(p);
    std::vector<float> fb, cb;
    dof_to_bricks<float>(h->levels[fl], fine, fb);
    cb.assign(h->levels[fl + 1].nleaf() * 512, 0.f);
    restrict_to<float>(h->levels[fl], h->levels[fl + 1], fb.data(), cb.data());
    bricks_to_dof<float>(h->levels[fl + 1], cb, coarse);
    return 0;
}

int orc_prolong(void* p, int fl, const float* coarse, float* fine, float alpha) {
    Handle* h = static_cast<Handle*>(p);
    std::vector<float> fb, cb;
    dof_to_bricks<float>(h->levels[fl], fine, fb);
    dof_to_bricks<float>(h->levels[fl + 1], coarse, cb);
    prolong_add<float>(h->levels[fl], h->levels[fl + 1], cb.data(), fb.data(), alpha);
    bricks_to_dof<float>(h->levels[fl], fb, fine);
    return 0;
}

int orc_coarsest(void* p, const float* b, float* x) {
    Handle* h = static_cast<Handle*>(p);
    const Level& L = h->levels.back();
    std::vector<float> bb, xb;
    dof_to_bricks<float>(L, b, bb);
    xb.assign(bb.size(), 0.f);
    coarsest_solve<float>(h->K, bb.data(), xb.data());
    bricks_to_dof<float>(L, xb, x);
    return 0;
}

int orc_precondition(void* p, int smoother, const float* r, float* z) {
    Handle* h = static_cast<Handle*>(p);
    if (smoother == kSPAI)
        for (auto& L : h->levels) build_spai0(L);
    h->s32.smoother = smoother;
    const Level& L0 = h->levels[0];
    std::vector<float> rb, zb;
    dof_to_bricks<float>(L0, r, rb);
    zb.assign(rb.size(), 0.f);
    h->s32.precond(0, zb.data(), rb.data(), true);
    bricks_to_dof<float>(L0, zb, z);
    return 0;
}

static void spai_all(Handle* h, int smoother) {
    if (smoother == kSPAI)
        for (auto& L : h->levels) build_spai0(L);
}

int orc_solve_pcg(void* p, int smoother, float tol, int maxit, float* x, const float* b, int* iters, float* hist,
                  int hist_cap) {
    Handle* h = static_cast<Handle*>(p);
    spai_all(h, smoother);
    h->s32.smoother = smoother;
    const Level& L0 = h->levels[0];
    std::vector<float> xb, bb;
    dof_to_bricks<float>(L0, x, xb);
    dof_to_bricks<float>(L0, b, bb);
    std::vector<float> history;
    const int st = h->s32.solve_pcg(xb.data(), bb.data(), tol, maxit, iters, &history);
    bricks_to_dof<float>(L0, xb, x);
    for (int i = 0; hist && i < hist_cap && i < int(history.size()); ++i) hist[i] = history[i];
    return st;
}

int orc_solve_mg(void* p, int smoother, float tol, int maxit, float* x, const float* b, int* iters) {
    Handle* h = static_cast<Handle*>(p);
    spai_all(h, smoother);
    h->s32.smoother = smoother;
    const Level& L0 = h->levels[0];
    std::vector<float> xb, bb;
    dof_to_bricks<float>(L0, x, xb);
    dof_to_bricks<float>(L0, b, bb);
    const int st = h->s32.solve_mg(xb.data(), bb.data(), tol, maxit, iters);
    bricks_to_dof<float>(L0, xb, x);
    return st;
}

// FP64 twin of solveMultigridPCG on the same (float) coefficients
int orc_solve_pcg64(void* p, int smoother, double tol, int maxit, double* x, const double* b, int* iters) {
    Handle* h = static_cast<Handle*>(p);
    spai_all(h, smoother);
    if (!h->s64_ready) {
        h->s64.init(h->levels, h->K);
        h->s64_ready = true;
    }
    h->s64.smoother = smoother;
    const Level& L0 = h->levels[0];
    std::vector<double> xb, bb;
    dof_to_bricks<double>(L0, x, xb);
    dof_to_bricks<double>(L0, b, bb);
    const int st = h->s64.solve_pcg(xb.data(), bb.data(), tol, maxit, iters, nullptr);
    bricks_to_dof<double>(L0, xb, x);
    return st;
}

// A x in FP64 on level l (true-residual checks)
int orc_apply64(void* p, int l, const double* x, double* y) {
    Handle* h = static_cast<Handle*>(p);
    const Level& L = h->levels[l];
    std::vector<double> xb, yb;
    dof_to_bricks<double>(L, x, xb);
    yb.assign(xb.size(), 0.0);
    apply<double>(L, kLap, 0.0, xb.data(), nullptr, yb.data());
    bricks_to_dof<double>(L, yb, y);
    return 0;
}

}  // extern "C"
