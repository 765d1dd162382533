// ORACLE - TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's variational viscosity solve
// (SIMD_UAAMG/simd_viscosity3d.{h,cpp}, volume_fractions.cpp,
// packed3grids.cpp). It is the checker for the HIP product: only tests, the
// smoke entry point and the benchmark's CPU-baseline leg may load it.
//
// Grids are the std::map leaves of vdb_grid.h in OpenVDB tree order. Each
// coefficient grid is built on the positions where the reference leaves the
// value active (the only positions whose value reaches a result: every other
// position multiplies a zero velocity); reads follow the reference's two
// conventions: the level operator reads a missing leaf as the level's
// default constant (light_weight_applier, :2991-2997), the coarsening reads
// it as 0 (pattern grids built by TopologyCopy with background 0, :1352-1411).
// Arithmetic is per value in the AVX code's order; fused multiply-adds only
// where the reference calls _mm256_fmadd_ps (fma() here; build with
// -ffp-contract=off).
//
// Restated OpenVDB behaviour (not in the reference tree; version unpinned):
// dilateActiveValues creates a neighbour leaf exactly when a dilated voxel
// lands in it; BoxSampler::sample is the trilinear a + T((b - a) * w) form.
// Sample positions of the subcell SDF are taken as exact half-integers, which
// they are for dyadic voxel sizes (dx = 2^-k); for other dx the reference's
// transform rounding is parity unpinned.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <unordered_map>
#include <vector>

#include "vdb_grid.h"

namespace orc {
namespace visc {

using CSet = std::set<Coord, TreeLess>;
using BGrid = Grid<uint8_t>;

inline Coord off3(Coord c, int a, int d) {
    c[a] += d;
    return c;
}
inline Coord plus(Coord a, Coord b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Coord unit(int a) {
    Coord c{0, 0, 0};
    c[a] = 1;
    return c;
}

// ------------------------------------------------------------ geometry
// volume_fractions.cpp:62-126 and util.h sort (ascending)
template <class T>
T sorted_tet_fraction(T p0, T p1, T p2, T p3) {
    return p0 * p0 * p0 / ((p0 - p1) * (p0 - p2) * (p0 - p3));
}
template <class T>
T sorted_prism_fraction(T p0, T p1, T p2, T p3) {
    const T a = p0 / (p0 - p2), b = p0 / (p0 - p3), c = p1 / (p1 - p3), d = p1 / (p1 - p2);
    return a * b * (1 - d) + b * (1 - c) * d + c * d;
}
float vf4(float p0, float p1, float p2, float p3) {
    auto sw = [](float& a, float& b) {
        if (a > b) std::swap(a, b);
    };
    sw(p0, p1);
    sw(p2, p3);
    sw(p0, p2);
    sw(p1, p3);
    sw(p1, p2);
    if (p3 <= 0) return 1;
    if (p2 <= 0) return 1 - sorted_tet_fraction(p3, p2, p1, p0);
    if (p1 <= 0) return sorted_prism_fraction(p0, p1, p2, p3);
    if (p0 <= 0) return sorted_tet_fraction(p0, p1, p2, p3);
    return 0;
}
float vf8(float p000, float p100, float p010, float p110, float p001, float p101, float p011, float p111) {
    if (p000 < 0.f && p100 < 0.f && p010 < 0.f && p110 < 0.f && p001 < 0.f && p101 < 0.f && p011 < 0.f &&
        p111 < 0.f)
        return 1.0f;
    return (vf4(p000, p001, p101, p011) + vf4(p000, p101, p100, p110) + vf4(p000, p010, p011, p110) +
            vf4(p101, p011, p111, p110) + 2 * vf4(p000, p011, p101, p110) + vf4(p100, p101, p001, p111) +
            vf4(p100, p001, p000, p010) + vf4(p100, p110, p111, p010) + vf4(p001, p111, p011, p010) +
            2 * vf4(p100, p111, p001, p010)) /
           12;
}

// OpenVDB BoxSampler::sample: floor, then trilinear a + T((b - a) * w)
float box_sample(const FGrid& g, double px, double py, double pz) {
    const int ix = int(std::floor(px)), iy = int(std::floor(py)), iz = int(std::floor(pz));
    const double u = px - ix, v = py - iy, w = pz - iz;
    float d[2][2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) d[i][j][k] = g.get({ix + i, iy + j, iz + k});
    auto lerp = [](float a, float b, double t) { return a + float(double(b - a) * t); };
    return lerp(lerp(lerp(d[0][0][0], d[0][0][1], w), lerp(d[0][1][0], d[0][1][1], w), v),
                lerp(lerp(d[1][0][0], d[1][0][1], w), lerp(d[1][1][0], d[1][1][1], w), v), u);
}

// leaves of a grid after dilateActiveValues(k iterations; 6 or 26 neighbours)
template <class T>
CSet dilated_leaves(const Grid<T>& g, int k, bool nn26) {
    CSet out;
    for (const auto& kv : g.leaves) {
        out.insert(kv.first);
        for (int off = 0; off < 512; ++off) {
            if (!kv.second.is_on(off)) continue;
            const Coord l = local_of(off);
            for (int ax = -1; ax <= 1; ++ax)
                for (int ay = -1; ay <= 1; ++ay)
                    for (int az = -1; az <= 1; ++az) {
                        if (!ax && !ay && !az) continue;
                        const int a[3] = {ax, ay, az};
                        int dm = 0, ds = 0;
                        for (int c = 0; c < 3; ++c) {
                            const int d = a[c] > 0 ? 8 - l[c] : (a[c] < 0 ? l[c] + 1 : 0);
                            dm = std::max(dm, d);
                            ds += d;
                        }
                        if ((nn26 ? dm : ds) <= k)
                            out.insert({kv.first.x + 8 * ax, kv.first.y + 8 * ay, kv.first.z + 8 * az});
                    }
        }
    }
    return out;
}

template <class T>
Leaf<T>& touch(Grid<T>& g, Coord origin, T fill) {
    auto it = g.leaves.find(origin);
    if (it != g.leaves.end()) return it->second;
    Leaf<T>& l = g.leaves[origin];
    for (int i = 0; i < 512; ++i) l.v[i] = fill;
    return l;
}
template <class T>
void set_on(Grid<T>& g, Coord c, T v, T fill) {
    Leaf<T>& l = touch(g, leaf_of(c), fill);
    l.v[offset_of(c)] = v;
    l.set_on(offset_of(c), true);
}
template <class T>
T get_or(const Grid<T>& g, Coord c, T missing) {
    const Leaf<T>* l = g.probe(leaf_of(c));
    return l ? l->v[offset_of(c)] : missing;
}

// ---------------------------------------------------------------- level
enum { kDiag = 0, kInvDiag = 3, kNeg = 6, kCross = 15, kSpai0 = 27, kNumGrids = 30 };
inline int neg_id(int c, int axis) { return kNeg + 3 * c + axis; }
inline int cross_id(int i, int j, int k) { return kCross + 4 * i + 2 * j + k; }

struct VLevel {
    int level = 0;
    float dx = 0, dt = 0, rho = 0;
    float default_diag = 1.f, default_off = 1.f;
    int64_t ndof = 0;
    IGrid dof[3];              // DOF index per channel, on = DOF, background -1
    FGrid g[kNumGrids];        // coefficient grids, only active positions stored
    float dflt[kNumGrids];     // level operator's constant for a missing leaf
    bool spai0_built = false;
    bool trimmed = false;
    // finest-level geometry used by the right-hand side
    FGrid voxel_vol, face_vol[3], edge_vol[3], mu;
    BGrid is_solid[3];
    CSet dof_topology;         // leaves of the DOF grids before pruning (finest)
    // linearised form for the level operator
    std::vector<Coord> org;                     // DOF leaves of any channel first, then coefficient-only leaves
    size_t nleaf_dof = 0;
    std::vector<std::array<int32_t, 27>> nb;    // 3x3x3 leaf neighbourhood, -1 = none
    std::vector<std::array<const float*, kNumGrids>> gp;
    std::vector<std::array<int32_t, 3 * 512>> idx;   // DOF index per (channel, voxel) of DOF leaves, -1 = none

    bool is_dof(int c, Coord p) const { return dof[c].is_on(p); }
    int32_t dof_at(int c, Coord p) const { return dof[c].get(p); }
};

void set_defaults(VLevel& L) {
    const float d = L.default_diag, o = L.default_off;
    for (int c = 0; c < 3; ++c) {
        L.dflt[kDiag + c] = d;
        L.dflt[kInvDiag + c] = 1.0f / d;
        for (int a = 0; a < 3; ++a) L.dflt[neg_id(c, a)] = a == c ? -2.0f * o : -o;
        L.dflt[kSpai0 + c] = 0.f;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) L.dflt[cross_id(i, j, k)] = (j + k) % 2 ? -o : o;
}

// set_dof (:504-537): prefix count over leaves in tree order, continuing the
// running DOF count across channels
void number_dofs(IGrid& g, int64_t& ndof) {
    for (auto& kv : g.leaves) {
        for (int off = 0; off < 512; ++off) {
            if (kv.second.is_on(off)) kv.second.v[off] = int32_t(ndof++);
            else kv.second.v[off] = -1;
        }
    }
}

void prune_empty(IGrid& g) {
    for (auto it = g.leaves.begin(); it != g.leaves.end();) {
        if (it->second.count() == 0) it = g.leaves.erase(it);
        else ++it;
    }
}

// ------------------------------------------------------ finest level
// L_with_level(viscosity, liquid, solid, dt, rho) (:22-50,258-963)
void build_finest(VLevel& L, const FGrid& liquid, const FGrid& solid, const FGrid& mu, float dt, float rho,
                  float dx) {
    L.level = 0;
    L.dx = dx;
    L.dt = dt;
    L.rho = rho;
    L.mu = mu;

    // subcell SDF and volumes (calculate_subcell_sdf_volume, :258-333)
    FGrid sub_sdf, sub_vol;
    sub_sdf.background = 1.5f * dx;
    sub_vol.background = 0.f;
    for (const Coord& o : dilated_leaves(liquid, 1, true))
        for (int i = 0; i <= 8; i += 8)
            for (int j = 0; j <= 8; j += 8)
                for (int k = 0; k <= 8; k += 8) {
                    const Coord so{2 * o.x + i, 2 * o.y + j, 2 * o.z + k};
                    touch(sub_sdf, so, sub_sdf.background);
                    touch(sub_vol, so, 0.f);
                }
    {
        std::vector<std::pair<Coord, Leaf<float>*>> lv;
        for (auto& kv : sub_sdf.leaves) lv.push_back({kv.first, &kv.second});
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t n = 0; n < int64_t(lv.size()); ++n) {
            for (int off = 0; off < 512; ++off) {
                const Coord l = local_of(off);
                const Coord s{lv[n].first.x + l.x, lv[n].first.y + l.y, lv[n].first.z + l.z};
                // subcell corner s sits at world (s/2 - 1/2) dx: liquid index (s - 1) / 2
                const float smp = box_sample(liquid, 0.5 * (s.x - 1), 0.5 * (s.y - 1), 0.5 * (s.z - 1));
                if (smp < sub_sdf.background) {
                    lv[n].second->v[off] = smp;
                    lv[n].second->set_on(off, true);
                }
            }
        }
    }
    {
        std::vector<std::pair<Coord, Leaf<float>*>> lv;
        for (auto& kv : sub_vol.leaves) lv.push_back({kv.first, &kv.second});
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t n = 0; n < int64_t(lv.size()); ++n) {
            for (int off = 0; off < 512; ++off) {
                const Coord l = local_of(off);
                const Coord s{lv[n].first.x + l.x, lv[n].first.y + l.y, lv[n].first.z + l.z};
                float p[2][2][2];
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j)
                        for (int k = 0; k < 2; ++k) p[i][j][k] = sub_sdf.get({s.x + i, s.y + j, s.z + k});
                const float f = vf8(p[0][0][0], p[1][0][0], p[0][1][0], p[1][1][0], p[0][0][1], p[1][0][1],
                                    p[0][1][1], p[1][1][1]);
                if (f > 0) {
                    lv[n].second->v[off] = f;
                    lv[n].second->set_on(off, true);
                }
            }
        }
    }

    // voxel / face / edge volumes (calculate_matrix_coefficients_volume, :335-418)
    const CSet vol_leaves = dilated_leaves(liquid, 2, true);
    for (const Coord& o : vol_leaves) {
        touch(L.voxel_vol, o, 0.f);
        for (int c = 0; c < 3; ++c) {
            touch(L.face_vol[c], o, 0.f);
            touch(L.edge_vol[c], o, 0.f);
        }
    }
    {
        std::vector<Coord> lv(vol_leaves.begin(), vol_leaves.end());
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t n = 0; n < int64_t(lv.size()); ++n) {
            Leaf<float>* vl = L.voxel_vol.probe(lv[n]);
            Leaf<float>* fl[3] = {L.face_vol[0].probe(lv[n]), L.face_vol[1].probe(lv[n]), L.face_vol[2].probe(lv[n])};
            Leaf<float>* el[3] = {L.edge_vol[0].probe(lv[n]), L.edge_vol[1].probe(lv[n]), L.edge_vol[2].probe(lv[n])};
            for (int off = 0; off < 512; ++off) {
                const Coord l = local_of(off);
                const Coord v{lv[n].x + l.x, lv[n].y + l.y, lv[n].z + l.z};
                float vc = 0, vfc[3] = {0, 0, 0}, vec[3] = {0, 0, 0};
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j)
                        for (int k = 0; k < 2; ++k) {
                            const Coord a{2 * v.x + i, 2 * v.y + j, 2 * v.z + k};
                            vc += sub_vol.get(a);
                            for (int c = 0; c < 3; ++c) {
                                Coord f = a;
                                f[c]--;
                                vfc[c] += sub_vol.get(f);
                                Coord e{a.x - 1, a.y - 1, a.z - 1};
                                e[c]++;
                                vec[c] += sub_vol.get(e);
                            }
                        }
                if (vc > 0) {
                    vl->v[off] = vc * 0.125f;
                    vl->set_on(off, true);
                }
                for (int c = 0; c < 3; ++c) {
                    if (vfc[c] > 0) {
                        fl[c]->v[off] = vfc[c] * 0.125f;
                        fl[c]->set_on(off, true);
                    }
                    if (vec[c] > 0) {
                        el[c]->v[off] = vec[c] * 0.125f;
                        el[c]->set_on(off, true);
                    }
                }
            }
        }
    }

    // DOF marking (mark_dof, :420-499)
    const CSet dof_leaves = dilated_leaves(liquid, 1, false);
    L.dof_topology = dof_leaves;
    for (int i = 0; i < 3; ++i) {
        // solid faces: face centre of channel i at index v lies at solid
        // (vertex-lattice) index v + 1/2 on the two other axes
        for (const Coord& o : dof_leaves) {
            Leaf<uint8_t>& sl = touch(L.is_solid[i], o, uint8_t(0));
            for (int off = 0; off < 512; ++off) {
                const Coord l = local_of(off);
                const Coord v{o.x + l.x, o.y + l.y, o.z + l.z};
                double p[3] = {v.x + 0.5, v.y + 0.5, v.z + 0.5};
                p[i] = v[i];
                if (box_sample(solid, p[0], p[1], p[2]) < 0) {
                    sl.v[off] = 1;
                    sl.set_on(off, true);
                }
            }
        }
        const int d1 = (i + 1) % 3, d2 = (i + 2) % 3;
        for (const Coord& o : dof_leaves) {
            Leaf<int32_t>& dl = touch(L.dof[i], o, int32_t(-1));
            const Leaf<uint8_t>* sl = L.is_solid[i].probe(o);
            for (int off = 0; off < 512; ++off) {
                const Coord l = local_of(off);
                const Coord v{o.x + l.x, o.y + l.y, o.z + l.z};
                bool on;
                if (sl && sl->is_on(off)) on = false;
                else if (L.face_vol[i].get(v) > 0) on = true;
                else if (L.voxel_vol.get(v) > 0 || L.voxel_vol.get(off3(v, i, -1)) > 0) on = true;
                else
                    on = L.edge_vol[d1].get(v) > 0 || L.edge_vol[d1].get(plus(v, unit(d2))) > 0 ||
                         L.edge_vol[d2].get(v) > 0 || L.edge_vol[d2].get(plus(v, unit(d1))) > 0;
                dl.set_on(off, on);
            }
        }
        L.dof[i].background = -1;
        number_dofs(L.dof[i], L.ndof);
        prune_empty(L.dof[i]);
    }

    // coefficients (calculate_matrix_coefficients + finest_level_matrix_setter, :539-908)
    const float factor = dt / (rho * dx * dx);
    L.default_off = factor * mu.background;
    L.default_diag = 1.0f + 8.0f * L.default_off;
    set_defaults(L);
    const float eps = 1e-1f;   // m_diag_frac_epsl
    for (int c = 0; c < 3; ++c) {
        const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        for (const auto& kv : L.dof[c].leaves) {
            for (int off = 0; off < 512; ++off) {
                if (!kv.second.is_on(off)) continue;
                const Coord l = local_of(off);
                const Coord v{kv.first.x + l.x, kv.first.y + l.y, kv.first.z + l.z};
                float diag = 0;
                float fv = L.face_vol[d0].get(v);
                if (fv < eps) fv = eps;
                diag += fv;
                {
                    const float t = 2 * factor * L.voxel_vol.get(v) * mu.get(v);
                    diag += t;
                }
                {
                    const Coord n = off3(v, d0, -1);
                    const float t = 2 * factor * L.voxel_vol.get(n) * mu.get(n);
                    diag += t;
                    if (L.is_dof(d0, n)) set_on(L.g[neg_id(c, d0)], v, -t, 0.f);
                }
                float fc[3] = {float(v.x), float(v.y), float(v.z)};
                fc[d0] -= 0.5f;
                auto mu_at = [&](int a, float da, int b, float db) {
                    float p[3] = {fc[0], fc[1], fc[2]};
                    p[a] += da;
                    if (b >= 0) p[b] += db;
                    return box_sample(mu, p[0], p[1], p[2]);
                };
                {   // + d1
                    const float ev = L.edge_vol[d2].get(plus(v, unit(d1)));
                    const float t = factor * ev * mu_at(d1, 0.5f, -1, 0.f);
                    diag += t;
                }
                {   // - d1
                    const float ev = L.edge_vol[d2].get(v);
                    const float t = factor * ev * mu_at(d1, -0.5f, -1, 0.f);
                    diag += t;
                    if (L.is_dof(d0, off3(v, d1, -1))) set_on(L.g[neg_id(c, d1)], v, -t, 0.f);
                }
                {   // + d2
                    const float ev = L.edge_vol[d1].get(plus(v, unit(d2)));
                    const float t = factor * ev * mu_at(d2, 0.5f, -1, 0.f);
                    diag += t;
                }
                {   // - d2
                    const float ev = L.edge_vol[d1].get(v);
                    const float t = factor * ev * mu_at(d2, -0.5f, -1, 0.f);
                    diag += t;
                    if (L.is_dof(d0, off3(v, d2, -1))) set_on(L.g[neg_id(c, d2)], v, -t, 0.f);
                }
                set_on(L.g[kDiag + c], v, diag, 0.f);
                set_on(L.g[kInvDiag + c], v, diag != 0 ? 1.0f / diag : 0.f, 0.f);
            }
        }
    }
    // cross terms of normal direction d0: active where both faces are DOFs
    for (int d0 = 0; d0 < 3; ++d0) {
        const int d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
        CSet cand;
        for (int ch : {d1, d2})
            for (const auto& kv : L.dof[ch].leaves)
                for (int ax = -1; ax <= 0; ++ax)
                    for (int ay = -1; ay <= 0; ++ay)
                        for (int az = -1; az <= 0; ++az)
                            cand.insert({kv.first.x + 8 * ax, kv.first.y + 8 * ay, kv.first.z + 8 * az});
        for (const Coord& o : cand) {
            for (int off = 0; off < 512; ++off) {
                const Coord l = local_of(off);
                const Coord v{o.x + l.x, o.y + l.y, o.z + l.z};
                const bool n1 = L.is_dof(d1, v), p1 = L.is_dof(d1, plus(v, unit(d1)));
                const bool n2 = L.is_dof(d2, v), p2 = L.is_dof(d2, plus(v, unit(d2)));
                auto term = [&](int j, int k) {
                    Coord e = v;
                    if (j) e = plus(e, unit(d1));
                    if (k) e = plus(e, unit(d2));
                    const float ev = L.edge_vol[d0].get(e);
                    float p[3] = {float(v.x), float(v.y), float(v.z)};
                    p[d1] += j ? 0.5f : -0.5f;
                    p[d2] += k ? 0.5f : -0.5f;
                    const float m = box_sample(mu, p[0], p[1], p[2]);
                    const float f = (j + k) % 2 ? -factor : factor;
                    return f * ev * m;
                };
                if (n1 && n2) set_on(L.g[cross_id(d0, 0, 0)], v, term(0, 0), 0.f);
                if (n1 && p2) set_on(L.g[cross_id(d0, 0, 1)], v, term(0, 1), 0.f);
                if (p1 && n2) set_on(L.g[cross_id(d0, 1, 0)], v, term(1, 0), 0.f);
                if (p1 && p2) set_on(L.g[cross_id(d0, 1, 1)], v, term(1, 1), 0.f);
            }
        }
    }
}

// --------------------------------------------------------- coarsening
// L_with_level(child, Coarsening) (:52-78): mark_dof_from_child (:1017-1058),
// coarse_matrix_coef_setter (:1061-1345)
void build_coarse(VLevel& C, const VLevel& F) {
    C.level = F.level + 1;
    C.dx = F.dx * 2.0f;
    C.dt = F.dt;
    C.rho = F.rho;
    C.default_diag = 1.0f;
    C.default_off = 1.0f;
    set_defaults(C);
    for (int i = 0; i < 3; ++i) {
        CSet coarse_leaves;
        for (const auto& kv : F.dof[i].leaves) coarse_leaves.insert(leaf_of({kv.first.x / 2, kv.first.y / 2, kv.first.z / 2}));
        for (const Coord& o : coarse_leaves) {
            Leaf<int32_t>& l = touch(C.dof[i], o, int32_t(-1));
            for (int off = 0; off < 512; ++off) {
                const Coord lc = local_of(off);
                const Coord c{o.x + lc.x, o.y + lc.y, o.z + lc.z};
                bool any = false;
                for (int a = 0; a < 2 && !any; ++a)
                    for (int b = 0; b < 2 && !any; ++b)
                        for (int d = 0; d < 2 && !any; ++d)
                            if (F.is_dof(i, {2 * c.x + a, 2 * c.y + b, 2 * c.z + d})) any = true;
                l.set_on(off, any);
            }
        }
        C.dof[i].background = -1;
        number_dofs(C.dof[i], C.ndof);
    }
    const float rap = 0.125f;
    for (int i = 0; i < 3; ++i) {
        const int d0 = i, d1 = (i + 1) % 3, d2 = (i + 2) % 3;
        const int dd[3] = {d0, d1, d2};
        for (const auto& kv : C.dof[i].leaves) {
            for (int off = 0; off < 512; ++off) {
                if (!kv.second.is_on(off)) continue;
                const Coord lc = local_of(off);
                const Coord c{kv.first.x + lc.x, kv.first.y + lc.y, kv.first.z + lc.z};
                const Coord fb{2 * c.x, 2 * c.y, 2 * c.z};
                float nd = 0.f;
                float nn[3] = {0.f, 0.f, 0.f};
                for (int ii = 0; ii < 2; ++ii)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int kk = 0; kk < 2; ++kk) {
                            Coord f = fb;
                            f[d0] += ii;
                            f[d1] += jj;
                            f[d2] += kk;
                            if (!F.is_dof(d0, f)) continue;
                            nd += get_or(F.g[kDiag + i], f, 0.f);
                            const int loc[3] = {ii, jj, kk};
                            for (int nd_ = 0; nd_ < 3; ++nd_) {
                                const float t = get_or(F.g[neg_id(i, dd[nd_])], f, 0.f);
                                if (!F.is_dof(d0, off3(f, dd[nd_], -1))) continue;
                                if (loc[nd_] == 0) nn[nd_] += t;
                                else nd += t * 2.0f;
                            }
                        }
                set_on(C.g[kDiag + i], c, rap * nd, 0.f);
                set_on(C.g[kInvDiag + i], c, nd != 0 ? 1.0f / (rap * nd) : 0.f, 0.f);
                for (int nd_ = 0; nd_ < 3; ++nd_)
                    if (C.is_dof(d0, off3(c, dd[nd_], -1))) set_on(C.g[neg_id(i, dd[nd_])], c, rap * nn[nd_], 0.f);
            }
        }
        // cross terms of normal direction i
        CSet cand;
        for (int ch : {d1, d2})
            for (const auto& kv : C.dof[ch].leaves)
                for (int ax = -1; ax <= 0; ++ax)
                    for (int ay = -1; ay <= 0; ++ay)
                        for (int az = -1; az <= 0; ++az)
                            cand.insert({kv.first.x + 8 * ax, kv.first.y + 8 * ay, kv.first.z + 8 * az});
        const FGrid* fx[4] = {&F.g[cross_id(i, 0, 0)], &F.g[cross_id(i, 0, 1)], &F.g[cross_id(i, 1, 0)],
                              &F.g[cross_id(i, 1, 1)]};
        for (const Coord& o : cand) {
            for (int off = 0; off < 512; ++off) {
                const Coord lc = local_of(off);
                const Coord c{o.x + lc.x, o.y + lc.y, o.z + lc.z};
                const bool n1 = C.is_dof(d1, c), p1 = C.is_dof(d1, plus(c, unit(d1)));
                const bool n2 = C.is_dof(d2, c), p2 = C.is_dof(d2, plus(c, unit(d2)));
                const bool pair[4] = {n1 && n2, n1 && p2, p1 && n2, p1 && p2};
                const Coord fb{2 * c.x, 2 * c.y, 2 * c.z};
                for (int t = 0; t < 4; ++t) {
                    if (!pair[t]) continue;
                    float s = 0.f;
                    for (int ii = 0; ii < 2; ++ii)
                        for (int jj = 0; jj < 2; ++jj)
                            for (int kk = 0; kk < 2; ++kk) {
                                Coord f = fb;
                                f[d0] += ii;
                                f[d1] += jj;
                                f[d2] += kk;
                                switch (t) {
                                    case 0:
                                        s += get_or(*fx[0], f, 0.f);
                                        if (kk == 0) s += get_or(*fx[1], f, 0.f);
                                        if (jj == 0) s += get_or(*fx[2], f, 0.f);
                                        if (kk == 0 && jj == 0) s += get_or(*fx[3], f, 0.f);
                                        break;
                                    case 1:
                                        if (kk == 1) s += get_or(*fx[1], f, 0.f);
                                        if (jj == 0 && kk == 1) s += get_or(*fx[3], f, 0.f);
                                        break;
                                    case 2:
                                        if (jj == 1) s += get_or(*fx[2], f, 0.f);
                                        if (jj == 1 && kk == 0) s += get_or(*fx[3], f, 0.f);
                                        break;
                                    default:
                                        if (jj == 1 && kk == 1) s += get_or(*fx[3], f, 0.f);
                                }
                            }
                    set_on(C.g[cross_id(i, t >> 1, t & 1)], c, rap * s, 0.f);
                }
            }
        }
    }
}

// trim_default_nodes at level 0 (:1700-1722) with LaplacianWithLevel::
// trimDefaultNodes (simd_vdb_poisson.cpp:1770-1811): a leaf whose active
// values all lie within |eps| of the default becomes a tile of the default.
void trim_grid(FGrid& g, float dflt, float eps) {
    eps = std::abs(eps);
    for (auto it = g.leaves.begin(); it != g.leaves.end();) {
        float me = 0;
        for (int off = 0; off < 512; ++off)
            if (it->second.is_on(off)) me = std::max(me, std::abs(it->second.v[off] - dflt));
        if (me <= eps) it = g.leaves.erase(it);
        else ++it;
    }
}
void trim_level0(VLevel& L) {
    const float eps = 1e-3f;
    for (int c = 0; c < 3; ++c) {
        trim_grid(L.g[kDiag + c], L.default_diag, L.default_diag * eps);
        trim_grid(L.g[kInvDiag + c], 1.0f / L.default_diag, 1.0f / L.default_diag * eps);
        for (int a = 0; a < 3; ++a) trim_grid(L.g[neg_id(c, a)], L.dflt[neg_id(c, a)], L.dflt[neg_id(c, a)] * eps);
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                trim_grid(L.g[cross_id(c, j, k)], L.dflt[cross_id(c, j, k)], L.dflt[cross_id(c, j, k)] * eps);
    }
    L.trimmed = true;
}

// value the level operator reads: a missing leaf is the default constant
inline float opv(const VLevel& L, int gid, Coord p) { return get_or(L.g[gid], p, L.dflt[gid]); }

// build_SPAI0_matrix / SPAI0_collector (:1453-1698)
void build_spai0(VLevel& L) {
    if (L.spai0_built) return;
    L.spai0_built = true;
    for (int c = 0; c < 3; ++c) {
        const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const Coord e0 = unit(d0), e1 = unit(d1), e2 = unit(d2);
        FGrid& out = L.g[kSpai0 + c];
        for (const auto& kv : L.dof[c].leaves)
            for (int off = 0; off < 512; ++off) {
                if (!kv.second.is_on(off)) continue;
                const Coord lc = local_of(off);
                const Coord v{kv.first.x + lc.x, kv.first.y + lc.y, kv.first.z + lc.z};
                float den = 0.f, num = 0.f;
                const float dg = opv(L, kDiag + c, v);
                den += dg * dg;
                num = dg;
                for (int a = 0; a < 3; ++a) {
                    const Coord up = off3(v, a, 1), dn = off3(v, a, -1);
                    if (L.is_dof(d0, up)) {
                        const float t = opv(L, neg_id(c, a), up);
                        den += t * t;
                    }
                    if (L.is_dof(d0, dn)) {
                        const float t = opv(L, neg_id(c, a), v);
                        den += t * t;
                    }
                }
                auto sq = [&](bool on, int gid, Coord p) {
                    if (on) {
                        const float t = opv(L, gid, p);
                        den += t * t;
                    }
                };
                const Coord m0{v.x - e0.x, v.y - e0.y, v.z - e0.z};
                sq(L.is_dof(d1, v), cross_id(d2, 0, 0), v);
                sq(L.is_dof(d1, plus(v, e1)), cross_id(d2, 0, 1), v);
                sq(L.is_dof(d1, m0), cross_id(d2, 1, 0), m0);
                sq(L.is_dof(d1, plus(m0, e1)), cross_id(d2, 1, 1), m0);
                sq(L.is_dof(d2, v), cross_id(d1, 0, 0), v);
                sq(L.is_dof(d2, m0), cross_id(d1, 0, 1), m0);
                sq(L.is_dof(d2, plus(v, e2)), cross_id(d1, 1, 0), v);
                sq(L.is_dof(d2, plus(m0, e2)), cross_id(d1, 1, 1), m0);
                set_on(out, v, den != 0 ? num / den : 0.f, 0.f);
            }
    }
    // the level's linearised leaf pointers predate these grids
    for (size_t i = 0; i < L.gp.size(); ++i)
        for (int c = 0; c < 3; ++c) {
            const Leaf<float>* lf = L.g[kSpai0 + c].probe(L.org[i]);
            L.gp[i][kSpai0 + c] = lf ? lf->v : nullptr;
        }
}

// linearised leaves + neighbour table for the level operator
void linearise(VLevel& L) {
    CSet dl;
    for (int c = 0; c < 3; ++c)
        for (const auto& kv : L.dof[c].leaves) dl.insert(kv.first);
    CSet extra;
    for (int gid = 0; gid < kNumGrids; ++gid)
        for (const auto& kv : L.g[gid].leaves)
            if (!dl.count(kv.first)) extra.insert(kv.first);
    L.org.assign(dl.begin(), dl.end());
    L.nleaf_dof = L.org.size();
    L.org.insert(L.org.end(), extra.begin(), extra.end());
    std::map<Coord, int32_t, TreeLess> index;
    for (size_t i = 0; i < L.org.size(); ++i) index[L.org[i]] = int32_t(i);
    L.nb.assign(L.org.size(), {});
    L.gp.assign(L.org.size(), {});
    for (size_t i = 0; i < L.org.size(); ++i) {
        for (int n = 0; n < 27; ++n) {
            const Coord o{L.org[i].x + 8 * (n / 9 - 1), L.org[i].y + 8 * ((n / 3) % 3 - 1), L.org[i].z + 8 * (n % 3 - 1)};
            auto it = index.find(o);
            L.nb[i][n] = it == index.end() ? -1 : it->second;
        }
        for (int gid = 0; gid < kNumGrids; ++gid) {
            const Leaf<float>* lf = L.g[gid].probe(L.org[i]);
            L.gp[i][gid] = lf ? lf->v : nullptr;
        }
    }
    L.idx.assign(L.nleaf_dof, {});
    for (size_t i = 0; i < L.nleaf_dof; ++i)
        for (int c = 0; c < 3; ++c) {
            const Leaf<int32_t>* lf = L.dof[c].probe(L.org[i]);
            for (int off = 0; off < 512; ++off) L.idx[i][c * 512 + off] = lf && lf->is_on(off) ? lf->v[off] : -1;
        }
}

// ----------------------------------------------------- level operator
// Vectors are bricks [leaf][channel][512] over the DOF leaves; non-DOF
// entries stay 0 (every operation writes 0 there or leaves them).
enum Mode { kNormal = 0, kResidual = 1, kJacobi = 2, kSpai0Mode = 3, kRedGS = 4, kBlackGS = 5 };

template <class R>
struct Ctx {
    const VLevel& L;
    const R* x;
    // value of channel ch at coordinate (leaf i local l + d): 0 outside DOF leaves
    R xv(int ch, size_t i, Coord l) const {
        int a[3];
        for (int c = 0; c < 3; ++c) a[c] = l[c] < 0 ? 0 : (l[c] > 7 ? 2 : 1);
        const int32_t n = L.nb[i][a[0] * 9 + a[1] * 3 + a[2]];
        if (n < 0 || size_t(n) >= L.nleaf_dof) return R(0);
        const int off = ((l.x & 7) << 6) | ((l.y & 7) << 3) | (l.z & 7);
        return x[(size_t(n) * 3 + ch) * 512 + off];
    }
    float cv(int gid, size_t i, Coord l) const {
        int a[3];
        for (int c = 0; c < 3; ++c) a[c] = l[c] < 0 ? 0 : (l[c] > 7 ? 2 : 1);
        const int32_t n = L.nb[i][a[0] * 9 + a[1] * 3 + a[2]];
        if (n < 0 || !L.gp[n][gid]) return L.dflt[gid];
        const int off = ((l.x & 7) << 6) | ((l.y & 7) << 3) | (l.z & 7);
        return L.gp[n][gid][off];
    }
};

template <class R>
inline R fmaR(R a, R b, R c) {
    return std::fma(a, b, c);
}

// light_weight_applier::component_operator<c> (:2841-3582) for one voxel;
// returns (A x)_offdiag and leaves diag in *dg
template <class R>
R offdiag(const Ctx<R>& X, int c, size_t i, Coord l) {
    const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
    auto sh = [&](Coord p, int a, int d) { return off3(p, a, d); };
    const R xp = X.xv(c, i, sh(l, 0, 1)) * R(X.cv(neg_id(c, 0), i, sh(l, 0, 1)));
    const R xm = X.xv(c, i, sh(l, 0, -1)) * R(X.cv(neg_id(c, 0), i, l));
    const R yp = X.xv(c, i, sh(l, 1, 1)) * R(X.cv(neg_id(c, 1), i, sh(l, 1, 1)));
    const R ym = X.xv(c, i, sh(l, 1, -1)) * R(X.cv(neg_id(c, 1), i, l));
    const R zp = X.xv(c, i, sh(l, 2, 1)) * R(X.cv(neg_id(c, 2), i, sh(l, 2, 1)));
    const R zm = X.xv(c, i, sh(l, 2, -1)) * R(X.cv(neg_id(c, 2), i, l));
    R s = (xp + xm) + (yp + ym);
    s = s + (zp + zm);
    const Coord m0 = sh(l, d0, -1);
    // normal d2: couples with channel d1
    const R t0 = X.xv(d1, i, l) * R(X.cv(cross_id(d2, 0, 0), i, l));
    const R t1 = R(X.cv(cross_id(d2, 0, 1), i, l)) * X.xv(d1, i, sh(l, d1, 1));
    const R t2 = R(X.cv(cross_id(d2, 1, 0), i, m0)) * X.xv(d1, i, m0);
    const R t3 = R(X.cv(cross_id(d2, 1, 1), i, m0)) * X.xv(d1, i, sh(m0, d1, 1));
    // normal d1: couples with channel d2
    const R u0 = R(X.cv(cross_id(d1, 0, 0), i, l)) * X.xv(d2, i, l);
    const R u1 = R(X.cv(cross_id(d1, 0, 1), i, m0)) * X.xv(d2, i, m0);
    const R u2 = R(X.cv(cross_id(d1, 1, 0), i, l)) * X.xv(d2, i, sh(l, d2, 1));
    const R u3 = R(X.cv(cross_id(d1, 1, 1), i, m0)) * X.xv(d2, i, sh(m0, d2, 1));
    const R T2 = (t0 + t1) + (t2 + t3);
    const R T1 = (u0 + u1) + (u2 + u3);
    s = s + T2;
    s = s + T1;
    return s;
}

// y = op(x, b) on channel c (y may alias x only for the Gauss-Seidel modes)
template <class R>
void apply_channel(const VLevel& L, int c, int mode, const R* x, const R* b, R* y) {
    Ctx<R> X{L, x};
    const R w = R(1.0f);   // m_damped_jacobi_coef
    const int64_t n = int64_t(L.nleaf_dof);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; ++i) {
        const int32_t* id = &L.idx[i][c * 512];
        for (int row = 0; row < 64; ++row) {
            bool any = false;
            for (int z = 0; z < 8; ++z) any |= id[row * 8 + z] >= 0;
            if (!any) continue;   // rows without a DOF are not written
            R res[8];
            for (int z = 0; z < 8; ++z) {
                const int off = row * 8 + z;
                const Coord l = local_of(off);
                const size_t k = (size_t(i) * 3 + c) * 512 + off;
                const R s = offdiag(X, c, size_t(i), l);
                const R lhs = x[k];
                const R dg = R(X.cv(kDiag + c, size_t(i), l));
                R r;
                switch (mode) {
                    case kRedGS:
                    case kBlackGS: {
                        R t = b[k] - s;
                        t = t * R(X.cv(kInvDiag + c, size_t(i), l));
                        t = t * w;
                        r = fmaR(lhs, R(1.0f) - w, t);
                        const bool red = ((l.x + l.y + l.z) & 1) == 0;
                        if (red != (mode == kRedGS)) r = lhs;
                        break;
                    }
                    case kJacobi: {
                        R a = fmaR(lhs, dg, s);
                        R t = b[k] - a;
                        t = t * R(X.cv(kInvDiag + c, size_t(i), l));
                        r = fmaR(w, t, lhs);
                        break;
                    }
                    case kSpai0Mode: {
                        R a = fmaR(lhs, dg, s);
                        R t = b[k] - a;
                        t = t * R(X.cv(kSpai0 + c, size_t(i), l));
                        r = fmaR(w, t, lhs);
                        break;
                    }
                    case kResidual:
                        r = b[k] - fmaR(lhs, dg, s);
                        break;
                    default:
                        r = fmaR(lhs, dg, s);
                }
                res[z] = id[off] >= 0 ? r : R(0);
            }
            // Gauss-Seidel passes only read the other colour of this channel,
            // so writing in place is order independent
            for (int z = 0; z < 8; ++z) y[(size_t(i) * 3 + c) * 512 + row * 8 + z] = res[z];
        }
    }
}

template <class R>
void apply_all(const VLevel& L, int mode, const R* x, const R* b, R* y) {
    for (int c = 0; c < 3; ++c) apply_channel(L, c, mode, x, b, y);
}
// XYZ_RBGS_apply / ZYX_RBGS_apply (:2221-2302), in place on x
template <class R>
void rbgs(const VLevel& L, bool xyz, R* x, const R* b) {
    if (xyz) {
        for (int c = 0; c < 3; ++c) {
            apply_channel(L, c, kRedGS, x, b, x);
            apply_channel(L, c, kBlackGS, x, b, x);
        }
    } else {
        for (int c = 2; c >= 0; --c) {
            apply_channel(L, c, kBlackGS, x, b, x);
            apply_channel(L, c, kRedGS, x, b, x);
        }
    }
}

// restriction (:2353-2413): coarse = 1/8 sum of on fine children
template <class R>
void restrict_to(const VLevel& F, const VLevel& C, const R* fine, R* coarse) {
    std::map<Coord, int64_t, TreeLess> fidx;
    for (size_t i = 0; i < F.nleaf_dof; ++i) fidx[F.org[i]] = int64_t(i);
    for (size_t i = 0; i < C.nleaf_dof; ++i)
        for (int c = 0; c < 3; ++c)
            for (int off = 0; off < 512; ++off) {
                if (C.idx[i][c * 512 + off] < 0) continue;
                const Coord lc = local_of(off);
                const Coord g{C.org[i].x + lc.x, C.org[i].y + lc.y, C.org[i].z + lc.z};
                R s = 0;
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                        for (int d = 0; d < 2; ++d) {
                            const Coord f{2 * g.x + a, 2 * g.y + b, 2 * g.z + d};
                            if (!F.is_dof(c, f)) continue;
                            s += fine[(size_t(fidx[leaf_of(f)]) * 3 + c) * 512 + offset_of(f)];
                        }
                coarse[(i * 3 + c) * 512 + off] = s * R(0.125f) * R(0.5f) * R(2.0f);
            }
}
// prolongation<true> (:2659-2729): fine += coarse on on fine voxels
template <class R>
void prolong_add(const VLevel& F, const VLevel& C, const R* coarse, R* fine) {
    std::map<Coord, int64_t, TreeLess> cidx;
    for (size_t i = 0; i < C.nleaf_dof; ++i) cidx[C.org[i]] = int64_t(i);
    for (size_t i = 0; i < F.nleaf_dof; ++i)
        for (int c = 0; c < 3; ++c)
            for (int off = 0; off < 512; ++off) {
                if (F.idx[i][c * 512 + off] < 0) continue;
                const Coord lf = local_of(off);
                const Coord f{F.org[i].x + lf.x, F.org[i].y + lf.y, F.org[i].z + lf.z};
                const Coord g{f.x >> 1, f.y >> 1, f.z >> 1};
                if (!C.is_dof(c, g)) continue;
                const R cv = coarse[(size_t(cidx[leaf_of(g)]) * 3 + c) * 512 + offset_of(g)];
                fine[(i * 3 + c) * 512 + off] += cv;
            }
}

// --------------------------------------------------- explicit matrix
// explicit_matrix_reducer (:2415-2596), rows sorted by column as Eigen's
// setFromTriplets leaves them
struct CSR {
    int64_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> col;
    std::vector<float> val;
};

void explicit_matrix(const VLevel& L, CSR& M) {
    std::vector<std::vector<std::pair<int32_t, float>>> rows(L.ndof);
    for (int c = 0; c < 3; ++c) {
        const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const Coord e1 = unit(d1), e2 = unit(d2);
        for (const auto& kv : L.dof[c].leaves)
            for (int off = 0; off < 512; ++off) {
                if (!kv.second.is_on(off)) continue;
                const Coord lc = local_of(off);
                const Coord v{kv.first.x + lc.x, kv.first.y + lc.y, kv.first.z + lc.z};
                auto& row = rows[kv.second.v[off]];
                row.push_back({kv.second.v[off], opv(L, kDiag + c, v)});
                for (int f = 0; f < 6; ++f) {
                    const int a = f / 2;
                    const Coord nb = off3(v, a, f % 2 == 0 ? 1 : -1);
                    const int32_t nd = L.dof_at(d0, nb);
                    if (nd == -1) continue;
                    row.push_back({nd, f % 2 == 0 ? opv(L, neg_id(c, a), nb) : opv(L, neg_id(c, a), v)});
                }
                const Coord m0 = off3(v, d0, -1);
                auto add = [&](int ch, Coord p, int gid, Coord q) {
                    const int32_t nd = L.dof_at(ch, p);
                    if (nd != -1) row.push_back({nd, opv(L, gid, q)});
                };
                add(d1, v, cross_id(d2, 0, 0), v);
                add(d1, plus(v, e1), cross_id(d2, 0, 1), v);
                add(d1, m0, cross_id(d2, 1, 0), m0);
                add(d1, plus(m0, e1), cross_id(d2, 1, 1), m0);
                add(d2, v, cross_id(d1, 0, 0), v);
                add(d2, m0, cross_id(d1, 0, 1), m0);
                add(d2, plus(v, e2), cross_id(d1, 1, 0), v);
                add(d2, plus(m0, e2), cross_id(d1, 1, 1), m0);
            }
    }
    M.n = L.ndof;
    M.ptr.assign(M.n + 1, 0);
    M.col.clear();
    M.val.clear();
    for (int64_t r = 0; r < M.n; ++r) {
        auto& row = rows[r];
        std::stable_sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < row.size(); ++k) {
            if (k > 0 && row[k].first == row[k - 1].first) M.val.back() += row[k].second;
            else {
                M.col.push_back(row[k].first);
                M.val.push_back(row[k].second);
            }
        }
        M.ptr[r + 1] = int64_t(M.col.size());
    }
}

// Eigen::ConjugateGradient<RowMajor, Upper|Lower> with its diagonal
// preconditioner, zero guess, 10 iterations, tolerance = epsilon (:2598-2657).
// Eigen's SpMV sums each row in column order; its vectorised dot-product
// order is parity unpinned.
template <class R>
void eigen_cg(const CSR& M, const std::vector<R>& b, std::vector<R>& x) {
    const int64_t n = M.n;
    std::vector<R> invd(n, R(1)), r = b, p(n), z(n), q(n);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = M.ptr[i]; k < M.ptr[i + 1]; ++k)
            if (M.col[k] == i && M.val[k] != 0.f) invd[i] = R(1) / R(M.val[k]);
    x.assign(n, R(0));
    auto dotv = [&](const std::vector<R>& a, const std::vector<R>& c) {
        R s = 0;
        for (int64_t i = 0; i < n; ++i) s += a[i] * c[i];
        return s;
    };
    auto spmv = [&](const std::vector<R>& v, std::vector<R>& out) {
        for (int64_t i = 0; i < n; ++i) {
            R t = 0;
            for (int64_t k = M.ptr[i]; k < M.ptr[i + 1]; ++k) t += R(M.val[k]) * v[M.col[k]];
            out[i] = t;
        }
    };
    const R tol = std::is_same<R, float>::value ? R(FLT_EPSILON) : R(DBL_EPSILON);
    const R rhs2 = dotv(r, r);
    if (rhs2 == R(0)) return;
    const R thr = std::max(R(tol * tol * rhs2), std::is_same<R, float>::value ? R(FLT_MIN) : R(DBL_MIN));
    if (rhs2 < thr) return;
    for (int64_t i = 0; i < n; ++i) p[i] = invd[i] * r[i];
    R abs_new = dotv(r, p);
    for (int it = 0; it < 10; ++it) {
        spmv(p, q);
        const R alpha = abs_new / dotv(p, q);
        for (int64_t i = 0; i < n; ++i) {
            x[i] = x[i] + alpha * p[i];
            r[i] = r[i] - alpha * q[i];
        }
        if (dotv(r, r) < thr) break;
        for (int64_t i = 0; i < n; ++i) z[i] = invd[i] * r[i];
        const R abs_old = abs_new;
        abs_new = dotv(r, z);
        const R beta = abs_new / abs_old;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
}

// ------------------------------------------------------- vectors
template <class R>
void dof_to_bricks(const VLevel& L, const float* dv, std::vector<R>& br) {
    br.assign(L.nleaf_dof * 3 * 512, R(0));
    for (size_t i = 0; i < L.nleaf_dof; ++i)
        for (int k = 0; k < 3 * 512; ++k)
            if (L.idx[i][k] >= 0) br[i * 3 * 512 + k] = R(dv[L.idx[i][k]]);
}
template <class R, class S>
void bricks_to_dof(const VLevel& L, const R* br, S* dv) {
    for (size_t i = 0; i < L.nleaf_dof; ++i)
        for (int k = 0; k < 3 * 512; ++k)
            if (L.idx[i][k] >= 0) dv[L.idx[i][k]] = S(br[i * 3 * 512 + k]);
}

// lv_dot (:3756-3767): per channel grid_dot_op (float per leaf, leaves in
// tree order), channels accumulated in float
template <class R>
R lv_dot(const VLevel& L, const R* a, const R* b) {
    R t = 0;
    for (int c = 0; c < 3; ++c) {
        R s = 0;
        for (size_t i = 0; i < L.nleaf_dof; ++i) {
            R p = 0;
            for (int off = 0; off < 512; ++off)
                if (L.idx[i][c * 512 + off] >= 0) p += a[(i * 3 + c) * 512 + off] * b[(i * 3 + c) * 512 + off];
            s += p;
        }
        t += s;
    }
    return t;
}
// lv_abs_max (:3742-3754): NaN-propagating within a channel, channels joined
// with std::max(result, channel)
template <class R>
R lv_absmax(const VLevel& L, const R* a) {
    R res = 0;
    for (int c = 0; c < 3; ++c) {
        R m = 0;
        bool stop = false;
        for (size_t i = 0; i < L.nleaf_dof && !stop; ++i)
            for (int off = 0; off < 512; ++off) {
                if (L.idx[i][c * 512 + off] < 0) continue;
                const R v = a[(i * 3 + c) * 512 + off];
                if (!std::isfinite(v)) {
                    m = v;
                    stop = true;
                    break;
                }
                m = std::max(m, std::abs(v));
            }
        res = std::max(res, m);
    }
    return res;
}

// --------------------------------------------------------- solver
template <class R>
struct VSolver {
    std::vector<VLevel>* lv = nullptr;
    CSR coarse;
    std::vector<std::vector<R>> lhs, rhs, tmp;

    void init(std::vector<VLevel>& levels) {
        lv = &levels;
        explicit_matrix(levels.back(), coarse);
        const size_t n = levels.size();
        lhs.resize(n);
        rhs.resize(n);
        tmp.resize(n);
        for (size_t l = 0; l < n; ++l) {
            const size_t sz = levels[l].nleaf_dof * 3 * 512;
            lhs[l].assign(sz, R(0));
            rhs[l].assign(sz, R(0));
            tmp[l].assign(sz, R(0));
        }
    }
    void solve_exact(R* x, const R* b) {
        const VLevel& L = lv->back();
        std::vector<R> bd(L.ndof, R(0)), xd;
        bricks_to_dof(L, b, bd.data());
        eigen_cg(coarse, bd, xd);
        for (size_t i = 0; i < L.nleaf_dof; ++i)
            for (int k = 0; k < 3 * 512; ++k)
                if (L.idx[i][k] >= 0) x[i * 3 * 512 + k] = xd[L.idx[i][k]];
    }
    // simd_viscosity3d::mucycle (:3828-3884), n = 2, mu = 2
    void mucycle(R* x, const R* b, bool precon, size_t level) {
        std::vector<VLevel>& L = *lv;
        if (level == L.size() - 1) {
            solve_exact(x, b);
            return;
        }
        if (precon) std::fill(x, x + L[level].nleaf_dof * 3 * 512, R(0));
        for (int i = 0; i < 2; ++i) rbgs(L[level], true, x, b);
        apply_all(L[level], kResidual, x, b, tmp[level].data());
        restrict_to(L[level], L[level + 1], tmp[level].data(), rhs[level + 1].data());
        mucycle(lhs[level + 1].data(), rhs[level + 1].data(), true, level + 1);
        mucycle(lhs[level + 1].data(), rhs[level + 1].data(), false, level + 1);
        prolong_add(L[level], L[level + 1], lhs[level + 1].data(), x);
        for (int i = 0; i < 2; ++i) rbgs(L[level], false, x, b);
    }
    // simd_viscosity3d::pcg_solve (:3694-3740)
    int pcg(R* x, const R* b, R tol, int maxit, float* hist) {
        const VLevel& L0 = (*lv)[0];
        const size_t sz = L0.nleaf_dof * 3 * 512;
        std::vector<R> r(sz), p(sz, R(0)), z(sz, R(0));
        apply_all(L0, kResidual, x, b, r.data());
        R nu = lv_absmax(L0, r.data());
        const R numax = tol * nu;
        if (hist) hist[0] = float(nu);
        if (nu <= numax) return 0;
        mucycle(p.data(), r.data(), true, 0);
        R rho = lv_dot(L0, p.data(), r.data());
        int it = 0;
        for (; it < maxit; ++it) {
            apply_all(L0, kNormal, p.data(), p.data(), z.data());
            const R sigma = lv_dot(L0, p.data(), z.data());
            const R alpha = rho / sigma;
            for (size_t k = 0; k < sz; ++k)
                if (L0.idx[k / 1536][k % 1536] >= 0) r[k] += (-alpha) * z[k];
            nu = lv_absmax(L0, r.data());
            if (hist) hist[it + 1] = float(nu);
            if (nu <= numax || it == maxit - 1) {
                for (size_t k = 0; k < sz; ++k)
                    if (L0.idx[k / 1536][k % 1536] >= 0) x[k] += alpha * p[k];
                return it;
            }
            std::fill(z.begin(), z.end(), R(0));
            mucycle(z.data(), r.data(), true, 0);
            const R rho_new = lv_dot(L0, z.data(), r.data());
            const R beta = rho_new / rho;
            rho = rho_new;
            for (size_t k = 0; k < sz; ++k)
                if (L0.idx[k / 1536][k % 1536] >= 0) x[k] += alpha * p[k];
            for (size_t k = 0; k < sz; ++k)
                if (L0.idx[k / 1536][k % 1536] >= 0) p[k] = z[k] + beta * p[k];
        }
        return it;
    }
};

// ------------------------------------------------ right-hand side
// build_rhs + rhs_reducer (:1724-2087)
void build_rhs(const VLevel& L, const VGrid& u, const VGrid& us, float* out_dof) {
    const float eps = 1e-1f;
    for (int c = 0; c < 3; ++c)
        for (const auto& kv : L.dof[c].leaves)
            for (int off = 0; off < 512; ++off) {
                if (!kv.second.is_on(off)) continue;
                const Coord lc = local_of(off);
                const Coord v{kv.first.x + lc.x, kv.first.y + lc.y, kv.first.z + lc.z};
                float fv = L.face_vol[c].get(v);
                if (fv < eps) fv = eps;
                out_dof[kv.second.v[off]] = u.get(v).v[c] * fv;
            }
    // solid pattern: union of the solid-face grids dilated 3 times (6 neighbours)
    BGrid uni;
    for (int c = 0; c < 3; ++c)
        for (const auto& kv : L.is_solid[c].leaves) {
            Leaf<uint8_t>& l = touch(uni, kv.first, uint8_t(0));
            for (int w = 0; w < 8; ++w) l.on[w] |= kv.second.on[w];
        }
    const CSet pattern = dilated_leaves(uni, 3, false);
    const float dx = L.dx, factor = L.dt / (L.rho * dx * dx);
    std::vector<std::pair<int64_t, float>> contrib;
    auto solid = [&](int ch, Coord p) { return L.is_solid[ch].is_on(p); };
    auto usv = [&](Coord p, int ch) { return us.get(p).v[ch]; };
    for (int c = 0; c < 3; ++c) {
        const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const Coord e0 = unit(d0), e1 = unit(d1), e2 = unit(d2);
        for (const Coord& o : pattern)
            for (int off = 0; off < 512; ++off) {
                const Coord lc = local_of(off);
                const Coord v{o.x + lc.x, o.y + lc.y, o.z + lc.z};
                const int32_t me = L.dof_at(d0, v);
                float fc[3] = {float(v.x), float(v.y), float(v.z)};
                fc[d0] -= 0.5f;
                auto mus = [&](int a, float da) {
                    float p[3] = {fc[0], fc[1], fc[2]};
                    p[a] += da;
                    return box_sample(L.mu, p[0], p[1], p[2]);
                };
                if (me != -1) {
                    float t = 0;
                    {
                        const Coord n = plus(v, e0);
                        const float term = 2 * factor * L.voxel_vol.get(v) * L.mu.get(v);
                        if (L.dof_at(d0, n) == -1 && solid(d0, n)) t += term * usv(n, d0);
                    }
                    {
                        const Coord n = off3(v, d0, -1);
                        const float term = 2 * factor * L.voxel_vol.get(n) * L.mu.get(n);
                        if (L.dof_at(d0, n) == -1 && solid(d0, n)) t += term * usv(n, d0);
                    }
                    {
                        const Coord n = plus(v, e1);
                        const float term = factor * L.edge_vol[d2].get(n) * mus(d1, 0.5f);
                        if (L.dof_at(d0, n) == -1 && solid(d0, n)) t += term * usv(n, d0);
                    }
                    {
                        const Coord n = off3(v, d1, -1);
                        const float term = factor * L.edge_vol[d2].get(v) * mus(d1, -0.5f);
                        if (L.dof_at(d0, n) == -1 && solid(d0, n)) t += term * usv(n, d0);
                    }
                    {
                        const Coord n = plus(v, e2);
                        const float term = factor * L.edge_vol[d1].get(n) * mus(d2, 0.5f);
                        if (L.dof_at(d0, n) == -1 && solid(d0, n)) t += term * usv(n, d0);
                    }
                    {
                        const Coord n = off3(v, d2, -1);
                        const float term = factor * L.edge_vol[d1].get(v) * mus(d2, -0.5f);
                        if (L.dof_at(d0, n) == -1 && solid(d0, n)) t += term * usv(n, d0);
                    }
                    if (t != 0) contrib.push_back({me, t});
                }
                const int32_t n1 = L.dof_at(d1, v), p1 = L.dof_at(d1, plus(v, e1));
                const int32_t n2 = L.dof_at(d2, v), p2 = L.dof_at(d2, plus(v, e2));
                const bool sn1 = solid(d1, v), sp1 = solid(d1, plus(v, e1));
                const bool sn2 = solid(d2, v), sp2 = solid(d2, plus(v, e2));
                auto term = [&](int j, int k) {
                    Coord e = v;
                    if (j) e = plus(e, e1);
                    if (k) e = plus(e, e2);
                    float p[3] = {float(v.x), float(v.y), float(v.z)};
                    p[d1] += j ? 0.5f : -0.5f;
                    p[d2] += k ? 0.5f : -0.5f;
                    const float f = (j + k) % 2 ? -factor : factor;
                    return f * L.edge_vol[d0].get(e) * box_sample(L.mu, p[0], p[1], p[2]);
                };
                if (n1 != -1 || n2 != -1) {
                    const float t = term(0, 0);
                    if (n1 == -1 && sn1) contrib.push_back({n2, -t * usv(v, d1)});
                    if (n2 == -1 && sn2) contrib.push_back({n1, -t * usv(v, d2)});
                }
                if (n1 != -1 || p2 != -1) {
                    const float t = term(0, 1);
                    if (n1 == -1 && sn1) contrib.push_back({p2, -t * usv(v, d1)});
                    if (p2 == -1 && sp2) contrib.push_back({n1, -t * usv(plus(v, e2), d2)});
                }
                if (p1 != -1 || n2 != -1) {
                    const float t = term(1, 0);
                    if (p1 == -1 && sp1) contrib.push_back({n2, -t * usv(plus(v, e1), d1)});
                    if (n2 == -1 && sn2) contrib.push_back({p1, -t * usv(v, d2)});
                }
                if (p1 != -1 || p2 != -1) {
                    const float t = term(1, 1);
                    if (p1 == -1 && sp1) contrib.push_back({p2, -t * usv(plus(v, e1), d1)});
                    if (p2 == -1 && sp2) contrib.push_back({p1, -t * usv(plus(v, e2), d2)});
                }
            }
    }
    std::unordered_map<int64_t, float> acc;
    for (const auto& kc : contrib) {
        auto it = acc.find(kc.first);
        if (it != acc.end()) it->second += kc.second;
        else acc[kc.first] = kc.second;
    }
    for (int64_t d = 0; d < L.ndof; ++d) {
        auto it = acc.find(d);
        if (it != acc.end()) out_dof[d] = out_dof[d] + it->second;
    }
}

struct Handle {
    std::vector<VLevel> levels;
    VSolver<float> s32;
    VSolver<double> s64;
};

FGrid load_scalar(int64_t n, const int32_t* org, const uint64_t* act, const float* val, float bg) {
    FGrid g;
    g.background = bg;
    for (int64_t i = 0; i < n; ++i) {
        Leaf<float>& l = g.leaves[Coord{org[3 * i], org[3 * i + 1], org[3 * i + 2]}];
        for (int w = 0; w < 8; ++w) l.on[w] = act ? act[8 * i + w] : ~uint64_t(0);
        std::memcpy(l.v, val + 512 * i, 512 * sizeof(float));
    }
    return g;
}
VGrid load_vec3(int64_t n, const int32_t* org, const uint64_t* act, const float* val, const float* bg) {
    VGrid g;
    for (int c = 0; c < 3; ++c) g.background.v[c] = bg ? bg[c] : 0.f;
    for (int64_t i = 0; i < n; ++i) {
        Leaf<Vec3>& l = g.leaves[Coord{org[3 * i], org[3 * i + 1], org[3 * i + 2]}];
        for (int w = 0; w < 8; ++w) l.on[w] = act ? act[8 * i + w] : ~uint64_t(0);
        for (int k = 0; k < 512; ++k)
            for (int c = 0; c < 3; ++c) l.v[k].v[c] = val[(512 * i + k) * 3 + c];
    }
    return g;
}

}  // namespace visc
}  // namespace orc

using namespace orc;
using namespace orc::visc;

extern "C" {

float orc_volume_fraction8(const float* p) { return vf8(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]); }
float orc_volume_fraction4(float a, float b, float c, float d) { return vf4(a, b, c, d); }

// grids: liquid SDF (voxel centres), solid SDF (vertex lattice: index i at
// world (i - 1/2) dx), viscosity (voxel centres). trim = 0 keeps the finest
// coefficients untrimmed (for identity checks against the coarse levels).
void* orc_visc_create(int64_t nl, const int32_t* l_org, const uint64_t* l_act, const float* l_val, float l_bg,
                      int64_t ns, const int32_t* s_org, const uint64_t* s_act, const float* s_val, float s_bg,
                      int64_t nm, const int32_t* m_org, const uint64_t* m_act, const float* m_val, float m_bg,
                      float dt, float rho, float dx, int64_t max_coarsest, int trim) {
    auto* h = new Handle();
    const FGrid liquid = load_scalar(nl, l_org, l_act, l_val, l_bg);
    const FGrid solid = load_scalar(ns, s_org, s_act, s_val, s_bg);
    const FGrid mu = load_scalar(nm, m_org, m_act, m_val, m_bg);
    h->levels.reserve(32);
    h->levels.emplace_back();
    build_finest(h->levels.back(), liquid, solid, mu, dt, rho, dx);
    while (h->levels.back().ndof > max_coarsest) {
        h->levels.emplace_back();
        build_coarse(h->levels.back(), h->levels[h->levels.size() - 2]);
    }
    if (trim) trim_level0(h->levels[0]);
    for (auto& L : h->levels) linearise(L);
    h->s32.init(h->levels);
    h->s64.init(h->levels);
    return h;
}

void orc_visc_destroy(void* p) { delete static_cast<Handle*>(p); }
int orc_visc_nlevels(void* p) { return int(static_cast<Handle*>(p)->levels.size()); }
int64_t orc_visc_level_ndof(void* p, int l) { return static_cast<Handle*>(p)->levels[l].ndof; }
int64_t orc_visc_level_nleaf(void* p, int l) { return int64_t(static_cast<Handle*>(p)->levels[l].nleaf_dof); }
float orc_visc_level_default(void* p, int l, int which) {
    const VLevel& L = static_cast<Handle*>(p)->levels[l];
    return which == 0 ? L.default_diag : L.default_off;
}

// DOF leaves of any channel in tree order: origins [n][3], per-channel masks
// [n][3][8], per-channel first DOF index [n][3] (-1 when the channel has none)
void orc_visc_level_topology(void* p, int l, int32_t* origins, uint64_t* masks, int64_t* dofbase) {
    const VLevel& L = static_cast<Handle*>(p)->levels[l];
    for (size_t i = 0; i < L.nleaf_dof; ++i) {
        origins[3 * i] = L.org[i].x;
        origins[3 * i + 1] = L.org[i].y;
        origins[3 * i + 2] = L.org[i].z;
        for (int c = 0; c < 3; ++c) {
            const Leaf<int32_t>* lf = L.dof[c].probe(L.org[i]);
            int64_t first = -1;
            for (int w = 0; w < 8; ++w) masks[(i * 3 + c) * 8 + w] = lf ? lf->on[w] : 0;
            if (lf)
                for (int off = 0; off < 512; ++off)
                    if (lf->is_on(off)) {
                        first = lf->v[off];
                        break;
                    }
            dofbase[i * 3 + c] = first;
        }
    }
}

// value the level operator reads for grid `gid` at every voxel of the DOF
// leaves ([n][512]) and whether the leaf is stored (kept) or read as default
void orc_visc_level_coefficients(void* p, int l, int gid, float* out, uint8_t* kept) {
    VLevel& L = static_cast<Handle*>(p)->levels[l];
    if (gid >= kSpai0) build_spai0(L);
    for (size_t i = 0; i < L.nleaf_dof; ++i) {
        const Leaf<float>* lf = L.g[gid].probe(L.org[i]);
        kept[i] = lf ? 1 : 0;
        for (int off = 0; off < 512; ++off) out[i * 512 + off] = lf ? lf->v[off] : L.dflt[gid];
    }
}

// active positions of one coefficient grid as (x, y, z, value) rows; count first
int64_t orc_visc_level_active(void* p, int l, int gid, int32_t* xyz, float* val) {
    VLevel& L = static_cast<Handle*>(p)->levels[l];
    if (gid >= kSpai0) build_spai0(L);
    int64_t n = 0;
    for (const auto& kv : L.g[gid].leaves)
        for (int off = 0; off < 512; ++off) {
            if (!kv.second.is_on(off)) continue;
            if (xyz) {
                const Coord lc = local_of(off);
                xyz[3 * n] = kv.first.x + lc.x;
                xyz[3 * n + 1] = kv.first.y + lc.y;
                xyz[3 * n + 2] = kv.first.z + lc.z;
                val[n] = kv.second.v[off];
            }
            ++n;
        }
    return n;
}

// finest geometry at DOF positions of channel c: face volume, and per DOF the
// solid flag (always 0 for DOFs) is implied; exported for closed-form checks
void orc_visc_face_volume(void* p, int c, float* out_dof) {
    const VLevel& L = static_cast<Handle*>(p)->levels[0];
    for (const auto& kv : L.dof[c].leaves)
        for (int off = 0; off < 512; ++off)
            if (kv.second.is_on(off)) {
                const Coord lc = local_of(off);
                out_dof[kv.second.v[off]] = L.face_vol[c].get({kv.first.x + lc.x, kv.first.y + lc.y, kv.first.z + lc.z});
            }
}
// volume grids sampled at arbitrary voxels: kind 0 = voxel, 1..3 = face c, 4..6 = edge c
void orc_visc_volume_at(void* p, int kind, int64_t n, const int32_t* xyz, float* out) {
    const VLevel& L = static_cast<Handle*>(p)->levels[0];
    const FGrid& g = kind == 0 ? L.voxel_vol : (kind <= 3 ? L.face_vol[kind - 1] : L.edge_vol[kind - 4]);
    for (int64_t i = 0; i < n; ++i) out[i] = g.get({xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
}
// DOF index of channel c at arbitrary voxels (-1 = none)
void orc_visc_dof_at(void* p, int l, int c, int64_t n, const int32_t* xyz, int32_t* out) {
    const VLevel& L = static_cast<Handle*>(p)->levels[l];
    for (int64_t i = 0; i < n; ++i) out[i] = L.dof_at(c, {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
}

// explicit matrix of a level as CSR (count with null arrays first)
int64_t orc_visc_level_matrix(void* p, int l, int64_t* ptr, int32_t* col, float* val) {
    VLevel& L = static_cast<Handle*>(p)->levels[l];
    CSR M;
    explicit_matrix(L, M);
    if (ptr) {
        std::memcpy(ptr, M.ptr.data(), M.ptr.size() * sizeof(int64_t));
        std::memcpy(col, M.col.data(), M.col.size() * sizeof(int32_t));
        std::memcpy(val, M.val.data(), M.val.size() * sizeof(float));
    }
    return int64_t(M.col.size());
}

void orc_visc_build_rhs(void* p, int64_t nu, const int32_t* u_org, const uint64_t* u_act, const float* u_val,
                        const float* u_bg, int64_t ns, const int32_t* s_org, const uint64_t* s_act,
                        const float* s_val, const float* s_bg, float* out_dof) {
    const VLevel& L = static_cast<Handle*>(p)->levels[0];
    const VGrid u = load_vec3(nu, u_org, u_act, u_val, u_bg);
    const VGrid us = load_vec3(ns, s_org, s_act, s_val, s_bg);
    build_rhs(L, u, us, out_dof);
}

// one level operation on DOF vectors: modes 0..5 as the applier, 6 = XYZ
// RBGS, 7 = ZYX RBGS (in place on x, result copied to y)
int orc_visc_apply(void* p, int l, int mode, const float* x, const float* b, float* y) {
    VLevel& L = static_cast<Handle*>(p)->levels[l];
    if (mode == kSpai0Mode) build_spai0(L);
    std::vector<float> xb, bb, yb;
    dof_to_bricks(L, x, xb);
    dof_to_bricks(L, b, bb);
    if (mode == kRedGS || mode == kBlackGS) {
        for (int c = 0; c < 3; ++c) apply_channel(L, c, mode, xb.data(), bb.data(), xb.data());
        bricks_to_dof(L, xb.data(), y);
        return 0;
    }
    if (mode == 6 || mode == 7) {
        rbgs(L, mode == 6, xb.data(), bb.data());
        bricks_to_dof(L, xb.data(), y);
        return 0;
    }
    yb.assign(xb.size(), 0.f);
    apply_all(L, mode, xb.data(), bb.data(), yb.data());
    bricks_to_dof(L, yb.data(), y);
    return 0;
}

int orc_visc_restrict(void* p, int fl, const float* fine, float* coarse) {
    auto* h = static_cast<Handle*>(p);
    std::vector<float> fb, cb(h->levels[fl + 1].nleaf_dof * 3 * 512, 0.f);
    dof_to_bricks(h->levels[fl], fine, fb);
    restrict_to(h->levels[fl], h->levels[fl + 1], fb.data(), cb.data());
    bricks_to_dof(h->levels[fl + 1], cb.data(), coarse);
    return 0;
}
int orc_visc_prolong(void* p, int fl, const float* coarse, float* fine) {
    auto* h = static_cast<Handle*>(p);
    std::vector<float> fb, cb;
    dof_to_bricks(h->levels[fl], fine, fb);
    dof_to_bricks(h->levels[fl + 1], coarse, cb);
    prolong_add(h->levels[fl], h->levels[fl + 1], cb.data(), fb.data());
    bricks_to_dof(h->levels[fl], fb.data(), fine);
    return 0;
}
int orc_visc_coarsest(void* p, const float* b, float* x) {
    auto* h = static_cast<Handle*>(p);
    const VLevel& L = h->levels.back();
    std::vector<float> bd(b, b + L.ndof), xd;
    eigen_cg(h->s32.coarse, bd, xd);
    std::copy(xd.begin(), xd.end(), x);
    return 0;
}
int orc_visc_precondition(void* p, const float* r, float* z) {
    auto* h = static_cast<Handle*>(p);
    const VLevel& L = h->levels[0];
    std::vector<float> rb, zb(L.nleaf_dof * 3 * 512, 0.f);
    dof_to_bricks(L, r, rb);
    h->s32.mucycle(zb.data(), rb.data(), true, 0);
    bricks_to_dof(L, zb.data(), z);
    return 0;
}
// pcg_solve; returns m_iteration, hist[0..it+1] = ||r||_inf
int orc_visc_solve(void* p, float tol, int maxit, float* x, const float* b, float* hist) {
    auto* h = static_cast<Handle*>(p);
    const VLevel& L = h->levels[0];
    std::vector<float> xb, bb;
    dof_to_bricks(L, x, xb);
    dof_to_bricks(L, b, bb);
    const int it = h->s32.pcg(xb.data(), bb.data(), tol, maxit, hist);
    bricks_to_dof(L, xb.data(), x);
    return it;
}
int orc_visc_solve64(void* p, double tol, int maxit, double* x, const double* b) {
    auto* h = static_cast<Handle*>(p);
    const VLevel& L = h->levels[0];
    std::vector<double> xb(L.nleaf_dof * 3 * 512, 0.0), bb(xb.size(), 0.0);
    for (size_t i = 0; i < L.nleaf_dof; ++i)
        for (int k = 0; k < 3 * 512; ++k)
            if (L.idx[i][k] >= 0) {
                xb[i * 1536 + k] = x[L.idx[i][k]];
                bb[i * 1536 + k] = b[L.idx[i][k]];
            }
    const int it = h->s64.pcg(xb.data(), bb.data(), tol, maxit, nullptr);
    bricks_to_dof(L, xb.data(), x);
    return it;
}

}  // extern "C"
