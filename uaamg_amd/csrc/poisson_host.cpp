// Host construction of the UAAMG Poisson level hierarchy (see poisson_host.h).
#include "poisson_host.h"

#include "build_timing.h"
#include "finest_device.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <thread>

namespace uaamg {

int host_threads() {
    static int n = [] {
        const char* env = std::getenv("UAAMG_HOST_THREADS");
        if (env && std::atoi(env) > 0) return std::atoi(env);
        unsigned hc = std::thread::hardware_concurrency();
        return int(hc == 0 ? 1 : std::min(hc, 64u));
    }();
    return n;
}

bool make_host_grid(int64_t nleaf, const int32_t* origins, const uint64_t* active, const float* values, int ncomp,
                    const float* background, HostGrid& out, bool borrow) {
    if (nleaf < 0 || (nleaf > 0 && (!origins || !values)) || (ncomp != 1 && ncomp != 3)) return false;
    out = HostGrid();
    out.ncomp = ncomp;
    for (int c = 0; c < 3; ++c) out.background[c] = background ? background[std::min(c, ncomp - 1)] : 0.f;
    auto org = [&](int64_t i) { return Coord{origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]}; };
    bool sorted = true;
    for (int64_t i = 0; i < nleaf; ++i) {
        const int32_t* o = origins + 3 * i;
        if ((o[0] & 7) || (o[1] & 7) || (o[2] & 7)) return false;
        if (i > 0 && !tree_order_less(org(i - 1), org(i))) sorted = false;
    }
    std::vector<int64_t> order;
    if (!sorted) {
        order.resize(size_t(nleaf));
        for (int64_t i = 0; i < nleaf; ++i) order[size_t(i)] = i;
        std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return tree_order_less(org(a), org(b)); });
    }
    auto src = [&](int64_t k) { return sorted ? k : order[size_t(k)]; };
    out.origin.resize(size_t(nleaf));
    for (int64_t k = 0; k < nleaf; ++k) {
        out.origin[size_t(k)] = org(src(k));
        if (k > 0 && out.origin[size_t(k)] == out.origin[size_t(k - 1)]) return false;
    }
    out.active.assign(size_t(nleaf) * 8, ~uint64_t(0));
    if (active) {
        if (sorted) std::copy(active, active + 8 * nleaf, out.active.begin());
        else
            for (int64_t k = 0; k < nleaf; ++k)
                std::copy(active + 8 * src(k), active + 8 * src(k) + 8, out.active.begin() + 8 * k);
    }
    const size_t per = size_t(kLeafSize) * size_t(ncomp);
    if (sorted && borrow) {
        out.values = values;
    } else {
        out.owned.reset(new float[size_t(nleaf) * per]);
        float* dst = out.owned.get();
        parallel_for(nleaf, [&](int64_t b, int64_t e) {
            for (int64_t k = b; k < e; ++k)
                std::copy(values + size_t(src(k)) * per, values + size_t(src(k) + 1) * per, dst + size_t(k) * per);
        });
        out.values = dst;
    }
    out.lookup.build(out.origin);
    return true;
}

namespace {

// Accessor with a one-leaf cache, the analogue of an OpenVDB ValueAccessor.
struct GridAccessor {
    const HostGrid* g;
    Coord last{INT32_MIN, INT32_MIN, INT32_MIN};
    int32_t last_idx = -1;
    explicit GridAccessor(const HostGrid& grid) : g(&grid) {}
    float value(const Coord& v, int comp = 0) {
        Coord o = leaf_origin_of(v);
        if (!(o == last)) {
            last = o;
            last_idx = g->lookup.find(o);
        }
        if (last_idx < 0) return g->background[comp];
        int off = voxel_offset(v.x & 7, v.y & 7, v.z & 7);
        return g->values[(size_t(last_idx) * kLeafSize + off) * g->ncomp + comp];
    }
};

// DOF-mask accessor over a level's compact leaves.
struct MaskAccessor {
    const PoissonLevelHost* l;
    Coord last{INT32_MIN, INT32_MIN, INT32_MIN};
    int32_t last_idx = -1;
    explicit MaskAccessor(const PoissonLevelHost& lv) : l(&lv) {}
    int32_t leaf(const Coord& v) {
        Coord o = leaf_origin_of(v);
        if (!(o == last)) {
            last = o;
            last_idx = l->lookup.find(o);
        }
        return last_idx;
    }
    bool on(const Coord& v) {
        int32_t li = leaf(v);
        if (li < 0) return false;
        return mask_test(&l->mask[size_t(li) * 8], voxel_offset(v.x & 7, v.y & 7, v.z & 7));
    }
    // effective coefficient (default when the leaf is absent)
    float coef(int kind, const Coord& v) {
        int32_t li = leaf(v);
        if (li < 0) return l->default_coef(kind);
        return l->coef[kind][size_t(li) * kLeafSize + voxel_offset(v.x & 7, v.y & 7, v.z & 7)];
    }
};

// fraction_inside for a segment (levelset_util.cpp:4-13)
inline float fraction_inside2(float phi_left, float phi_right) {
    if (phi_left < 0 && phi_right < 0) return 1;
    if (phi_left < 0 && phi_right >= 0) return phi_left / (phi_left - phi_right);
    if (phi_left >= 0 && phi_right < 0) return phi_right / (phi_right - phi_left);
    return 0;
}


// trimDefaultNodes (simd_vdb_poisson.cpp:1773-1811): a leaf whose active
// values all lie within |eps| of the default is removed; afterwards every
// voxel of that leaf reads the default.
bool trim_leaf(const uint64_t* m, float* v, float default_val, float eps) {
    float max_err = 0;
    for (int off = 0; off < kLeafSize; ++off) {
        if (!mask_test(m, off)) continue;
        float err = std::fabs(v[off] - default_val);
        if (err > max_err) max_err = err;
    }
    if (max_err <= eps) {
        std::fill(v, v + kLeafSize, default_val);
        return true;
    }
    return false;
}

void trim_kind(PoissonLevelHost& l, int kind, float default_val, float eps) {
    eps = std::fabs(eps);
    const size_t n = l.origin.size();
    l.keep[kind].assign(n, 1);
    parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i)
            if (trim_leaf(&l.mask[i * 8], &l.coef[kind][size_t(i) * kLeafSize], default_val, eps)) l.keep[kind][i] = 0;
    });
}

// the four stencil kinds trimmed, then 1/diag (0 where diag == 0, the
// default where the diagonal leaf is trimmed or the voxel holds no DOF), in
// one pass over each leaf
void trim_level(PoissonLevelHost& l) {
    const float term = -l.dt / (l.dx * l.dx);
    const float diagonal_term = -6.0f * term;
    const float epsilon = 1e-5f;
    const float dflt[4] = {diagonal_term, term, term, term};
    const float eps[4] = {std::fabs(diagonal_term * epsilon), std::fabs(term * epsilon), std::fabs(term * epsilon),
                          std::fabs(term * epsilon)};
    const size_t n = l.origin.size();
    const int kinds[4] = {kCoefDiag, kCoefX, kCoefY, kCoefZ};
    for (int k : kinds) l.keep[k].assign(n, 1);
    l.invdiag.resize(n * kLeafSize);
    const float inv_dflt = l.default_invdiag();
    parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            const uint64_t* m = &l.mask[i * 8];
            for (int j = 0; j < 4; ++j)
                if (trim_leaf(m, &l.coef[kinds[j]][size_t(i) * kLeafSize], dflt[j], eps[j])) l.keep[kinds[j]][i] = 0;
            float* inv = &l.invdiag[size_t(i) * kLeafSize];
            const float* d = &l.coef[kCoefDiag][size_t(i) * kLeafSize];
            if (!l.keep[kCoefDiag][i]) {
                std::fill(inv, inv + kLeafSize, inv_dflt);
                continue;
            }
            for (int off = 0; off < kLeafSize; ++off) {
                if (!mask_test(m, off)) {
                    inv[off] = inv_dflt;
                    continue;
                }
                inv[off] = d[off] == 0.f ? 0.f : 1.0f / d[off];
            }
        }
    });
}

}  // namespace

void finalize_topology(PoissonLevelHost& l) {
    const size_t n = l.origin.size();
    l.lookup.build(l.origin);
    l.dofbase.resize(n);
    int64_t run = 0;
    for (size_t i = 0; i < n; ++i) {
        l.dofbase[i] = run;
        run += mask_count(&l.mask[i * 8]);
    }
    l.ndof = run;
    l.nbr.assign(n * 6, -1);
    parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            for (int f = 0; f < 6; ++f) {
                Coord o = l.origin[i];
                o[f >> 1] += (f & 1) ? 8 : -8;
                l.nbr[i * 6 + f] = l.lookup.find(o);
            }
        }
    });
}

float PoissonLevelHost::default_coef(int kind) const {
    switch (kind) {
        case kCoefDiag: return 6.f * term;
        case kCoefSpai0: return default_spai0_apply();
        default: return -term;
    }
}

bool build_finest_level(const HostGrid& phi, const HostGrid& weights, float dt, float dx, PoissonLevelHost& l,
                        int device, std::string* err) {
    if (phi.ncomp != 1 || weights.ncomp != 3 || !(dx > 0.f)) return false;
    l = PoissonLevelHost();
    l.level = 0;
    l.dt = dt;
    l.dx = dx;
    l.term = dt / (dx * dx);
    const float dt_over_dx2 = l.term;
    const size_t nphi = phi.nleaf();
    l.full_origin = phi.origin;

    // per phi leaf: DOF mask and the four coefficient bricks
    StageClock clock;
    // every element is written by the per-voxel stage (uninitialised until then)
    std::unique_ptr<uint64_t[]> dmask(new uint64_t[nphi * 8 + 1]);
    std::unique_ptr<float[]> cd(new float[nphi * kLeafSize + 1]), cxyz[3];
    for (int c = 0; c < 3; ++c) cxyz[c].reset(new float[nphi * kLeafSize + 1]);
    clock.lap("level 0: allocate");

    // the per-voxel stage on the device (finest_device.h) or on the host threads
    if (device >= 0) {
        std::string e;
        float* const out[3] = {cxyz[0].get(), cxyz[1].get(), cxyz[2].get()};
        if (finest_coefficients_device(phi, weights, dt_over_dx2, device, dmask.get(), cd.get(), out, e) != 0) {
            if (err) *err = e;
            return false;
        }
    }
    if (device < 0) parallel_for(int64_t(nphi), [&](int64_t b, int64_t e) {
        GridAccessor phi_axr[6] = {GridAccessor(phi), GridAccessor(phi), GridAccessor(phi),
                                   GridAccessor(phi), GridAccessor(phi), GridAccessor(phi)};
        GridAccessor w_self(weights), w_pos[3] = {GridAccessor(weights), GridAccessor(weights), GridAccessor(weights)};
        for (int64_t li = b; li < e; ++li) {
            const Coord o = phi.origin[li];
            const float* pv = phi.leaf_values(li);
            const uint64_t* am = &phi.active[li * 8];
            std::fill(&dmask[li * 8], &dmask[li * 8] + 8, uint64_t(0));
            float* dleaf = &cd[li * kLeafSize];
            float* tleaf[3] = {&cxyz[0][li * kLeafSize], &cxyz[1][li * kLeafSize], &cxyz[2][li * kLeafSize]};
            std::fill(dleaf, dleaf + kLeafSize, 6.f * dt_over_dx2);
            for (int c = 0; c < 3; ++c) std::fill(tleaf[c], tleaf[c] + kLeafSize, -dt_over_dx2);
            for (int off = 0; off < kLeafSize; ++off) {
                if (!mask_test(am, off)) continue;
                const float phi_here = pv[off];
                if (!(phi_here < 0.f)) continue;
                const Coord g{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                float diagonal = 0.f;
                float term_xyz[3] = {0.f, 0.f, 0.f};
                // faces in the order +x,-x,+y,-y,+z,-z
                for (int f = 0; f < 6; ++f) {
                    const int comp = f >> 1;
                    const bool positive = (f & 1) == 0;
                    Coord pc = g;
                    float weight;
                    if (positive) {
                        Coord wc = g;
                        wc[comp] += 1;
                        weight = w_pos[comp].value(wc, comp);
                        pc[comp] += 1;
                    } else {
                        weight = w_self.value(g, comp);
                        pc[comp] -= 1;
                    }
                    const float phi_other = phi_axr[f].value(pc);
                    const float term = weight * dt_over_dx2;
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
                    mask_set(&dmask[li * 8], off);
                    dleaf[off] = diagonal;
                    for (int c = 0; c < 3; ++c) tleaf[c][off] = term_xyz[c];
                }
            }
        }
    });

    clock.lap("level 0: per-voxel stage");
    // compact: keep leaves holding DOFs, in tree order
    std::vector<int64_t> keep_idx;
    keep_idx.reserve(nphi);
    for (size_t i = 0; i < nphi; ++i)
        if (mask_count(&dmask[i * 8]) > 0) keep_idx.push_back(int64_t(i));
    const size_t n = keep_idx.size();
    l.origin.resize(n);
    l.mask.resize(n * 8);
    for (int k = 0; k < kNumCoef; ++k) l.coef[k].clear();
    for (int k = 0; k < 4; ++k) l.coef[k].resize(n * kLeafSize);
    parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
        for (int64_t k = b; k < e; ++k) {
            const int64_t i = keep_idx[size_t(k)];
            l.origin[size_t(k)] = phi.origin[size_t(i)];
            std::copy(&dmask[i * 8], &dmask[i * 8] + 8, &l.mask[size_t(k) * 8]);
            std::copy(&cd[i * kLeafSize], &cd[i * kLeafSize] + kLeafSize, &l.coef[kCoefDiag][size_t(k) * kLeafSize]);
            for (int c = 0; c < 3; ++c)
                std::copy(&cxyz[c][i * kLeafSize], &cxyz[c][i * kLeafSize] + kLeafSize,
                          &l.coef[1 + c][size_t(k) * kLeafSize]);
        }
    });
    clock.lap("level 0: compact");
    finalize_topology(l);
    clock.lap("level 0: topology");
    trim_level(l);
    clock.lap("level 0: trim + invdiag");
    return true;
}

void build_coarse_topology(const PoissonLevelHost& fine, PoissonLevelHost& l) {
    l = PoissonLevelHost();
    l.level = fine.level + 1;
    l.dt = fine.dt;
    l.dx = 2.0f * fine.dx;
    l.term = l.dt / (l.dx * l.dx);

    auto parent_origin = [](const Coord& fo) {
        // TouchCoarseLeafReducer touches Coord(origin / 2); origins are even so
        // the truncating division is exact, and the touched leaf is the one
        // containing that coordinate.
        return leaf_origin_of(Coord{fo.x / 2, fo.y / 2, fo.z / 2});
    };
    auto unique_tree_sorted = [](std::vector<Coord>& v) {
        std::sort(v.begin(), v.end(), tree_order_less);
        v.erase(std::unique(v.begin(), v.end()), v.end());
    };
    l.full_origin.reserve(fine.full_origin.size() / 4 + 1);
    for (const Coord& fo : fine.full_origin) l.full_origin.push_back(parent_origin(fo));
    unique_tree_sorted(l.full_origin);
    for (const Coord& fo : fine.origin) l.origin.push_back(parent_origin(fo));
    unique_tree_sorted(l.origin);

    const size_t n = l.origin.size();
    l.child.assign(n * 8, -1);
    l.mask.assign(n * 8, 0);
    for (size_t i = 0; i < n; ++i) {
        for (int b = 0; b < 8; ++b) {
            Coord fo{2 * l.origin[i].x + ((b >> 2) & 1) * 8, 2 * l.origin[i].y + ((b >> 1) & 1) * 8,
                     2 * l.origin[i].z + (b & 1) * 8};
            l.child[i * 8 + b] = fine.lookup.find(fo);
        }
    }
    // coarse voxel is on iff any of its 2^3 fine voxels is on
    parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            for (int off = 0; off < kLeafSize; ++off) {
                const int cx = off >> 6, cy = (off >> 3) & 7, cz = off & 7;
                const int oct = ((cx >= 4) << 2) | ((cy >= 4) << 1) | (cz >= 4);
                const int32_t fl = l.child[i * 8 + oct];
                if (fl < 0) continue;
                const uint64_t* fm = &fine.mask[size_t(fl) * 8];
                const int fx = 2 * (cx & 3), fy = 2 * (cy & 3), fz = 2 * (cz & 3);
                bool any = false;
                for (int ii = 0; ii < 2 && !any; ++ii)
                    for (int jj = 0; jj < 2 && !any; ++jj)
                        for (int kk = 0; kk < 2 && !any; ++kk)
                            any = mask_test(fm, voxel_offset(fx + ii, fy + jj, fz + kk));
                if (any) mask_set(&l.mask[i * 8], off);
            }
        }
    });
    finalize_topology(l);
}

void build_coarse_level(const PoissonLevelHost& fine, PoissonLevelHost& l) {
    StageClock clock;
    build_coarse_topology(fine, l);
    const size_t n = l.nleaf();
    clock.lap("coarse level: topology");
    // Galerkin coefficients (initializeFromFineLevel, simd_vdb_poisson.cpp:392-467);
    // every element is written below
    for (int k = 0; k < 4; ++k) l.coef[k].resize(n * kLeafSize);
    const float dt_over_dx2 = l.term;
    parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
        // the 2^3 fine voxels of a coarse voxel lie in one child leaf (octant of
        // the voxel); a fine voxel's -x/-y/-z neighbour lies in that leaf or in
        // its face neighbour, so every test and read is a direct array access
        auto on_at = [&](int32_t leaf, int lx, int ly, int lz) {
            return leaf >= 0 && mask_test(&fine.mask[size_t(leaf) * 8], voxel_offset(lx, ly, lz));
        };
        for (int64_t i = b; i < e; ++i) {
            float* dl = &l.coef[kCoefDiag][size_t(i) * kLeafSize];
            float* xl = &l.coef[kCoefX][size_t(i) * kLeafSize];
            float* yl = &l.coef[kCoefY][size_t(i) * kLeafSize];
            float* zl = &l.coef[kCoefZ][size_t(i) * kLeafSize];
            std::fill(dl, dl + kLeafSize, 6.0f * dt_over_dx2);
            std::fill(xl, xl + kLeafSize, -dt_over_dx2);
            std::fill(yl, yl + kLeafSize, -dt_over_dx2);
            std::fill(zl, zl + kLeafSize, -dt_over_dx2);
            const uint64_t* m = &l.mask[i * 8];
            for (int off = 0; off < kLeafSize; ++off) {
                if (!mask_test(m, off)) continue;
                const int cx = off >> 6, cy = (off >> 3) & 7, cz = off & 7;
                const int32_t fl = l.child[size_t(i) * 8 + size_t(((cx >= 4) << 2) | ((cy >= 4) << 1) | (cz >= 4))];
                const int32_t nx = fl >= 0 ? fine.nbr[size_t(fl) * 6 + 0] : -1;
                const int32_t ny = fl >= 0 ? fine.nbr[size_t(fl) * 6 + 2] : -1;
                const int32_t nz = fl >= 0 ? fine.nbr[size_t(fl) * 6 + 4] : -1;
                const int fx = 2 * (cx & 3), fy = 2 * (cy & 3), fz = 2 * (cz & 3);
                float diag = 0, x = 0, y = 0, z = 0;
                for (int ii = 0; ii < 2; ++ii)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int kk = 0; kk < 2; ++kk) {
                            const int lx = fx + ii, ly = fy + jj, lz = fz + kk;
                            if (!on_at(fl, lx, ly, lz)) continue;
                            const size_t at = size_t(fl) * kLeafSize + size_t(voxel_offset(lx, ly, lz));
                            diag += fine.coef[kCoefDiag][at];
                            if (lx > 0 ? on_at(fl, lx - 1, ly, lz) : on_at(nx, 7, ly, lz)) {
                                if (ii == 0) x += fine.coef[kCoefX][at];
                                else diag += 2 * fine.coef[kCoefX][at];
                            }
                            if (ly > 0 ? on_at(fl, lx, ly - 1, lz) : on_at(ny, lx, 7, lz)) {
                                if (jj == 0) y += fine.coef[kCoefY][at];
                                else diag += 2 * fine.coef[kCoefY][at];
                            }
                            if (lz > 0 ? on_at(fl, lx, ly, lz - 1) : on_at(nz, lx, ly, 7)) {
                                if (kk == 0) z += fine.coef[kCoefZ][at];
                                else diag += 2 * fine.coef[kCoefZ][at];
                            }
                        }
                const float factor = 0.5f * (1.0f / 8.0f);
                dl[off] = diag * factor;
                xl[off] = x * factor;
                yl[off] = y * factor;
                zl[off] = z * factor;
            }
        }
    });
    clock.lap("coarse level: Galerkin");
    trim_level(l);
    clock.lap("coarse level: trim + invdiag");
}

bool build_hierarchy(const HostGrid& phi, const HostGrid& weights, float dt, float dx, int64_t max_coarsest_dof,
                     PoissonHierarchyHost& out, int device, std::string* err) {
    out = PoissonHierarchyHost();
    out.nleaf_phi = int64_t(phi.nleaf());
    out.levels.emplace_back();
    if (!build_finest_level(phi, weights, dt, dx, out.levels.back(), device, err)) return false;
    StageClock clock;
    while (out.levels.back().ndof > max_coarsest_dof) {
        PoissonLevelHost next;
        build_coarse_level(out.levels.back(), next);
        out.levels.push_back(std::move(next));
        clock.lap("coarse level");
    }
    return true;
}

void build_spai0(PoissonHierarchyHost& h) {
    if (h.spai0_ready) return;
    for (auto& l : h.levels) {
        const size_t n = l.nleaf();
        // default for comparison in trimming: 6 / (7 * background of the diagonal grid)
        const float bg = 6.f * l.term;
        const float trim_default = 6.0f / (7.f * bg);
        l.coef[kCoefSpai0].assign(n * kLeafSize, trim_default);
        parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
            MaskAccessor axr(l), naxr(l);
            auto sqr = [](float v) { return v * v; };
            for (int64_t i = b; i < e; ++i) {
                const uint64_t* m = &l.mask[i * 8];
                const Coord o = l.origin[i];
                float* s = &l.coef[kCoefSpai0][size_t(i) * kLeafSize];
                for (int off = 0; off < kLeafSize; ++off) {
                    if (!mask_test(m, off)) continue;
                    const Coord g{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    float row = 0;
                    for (int c = 0; c < 3; ++c) {
                        Coord up = g, dn = g;
                        up[c] += 1;
                        dn[c] -= 1;
                        if (naxr.on(up)) row += sqr(axr.coef(kCoefX + c, up));
                        if (naxr.on(dn)) row += sqr(axr.coef(kCoefX + c, g));
                    }
                    const float d = axr.coef(kCoefDiag, g);
                    s[off] = d == 0 ? 0.f : d / (sqr(d) + row);
                }
            }
        });
        trim_kind(l, kCoefSpai0, trim_default, trim_default * 1e-5f);
        // trimmed leaves read the apply operator's default instead
        const float apply_default = l.default_spai0_apply();
        for (size_t i = 0; i < n; ++i)
            if (!l.keep[kCoefSpai0][i])
                std::fill(&l.coef[kCoefSpai0][i * kLeafSize], &l.coef[kCoefSpai0][i * kLeafSize] + kLeafSize,
                          apply_default);
    }
    h.spai0_ready = true;
}

void level_triplets(const PoissonLevelHost& l, std::vector<int32_t>& rows, std::vector<int32_t>& cols,
                    std::vector<float>& vals) {
    rows.clear();
    cols.clear();
    vals.clear();
    auto dof_at = [&](const Coord& g) -> int64_t {
        const int32_t li = l.lookup.find(leaf_origin_of(g));
        if (li < 0) return -1;
        const int off = voxel_offset(g.x & 7, g.y & 7, g.z & 7);
        const uint64_t* m = &l.mask[size_t(li) * 8];
        if (!mask_test(m, off)) return -1;
        int64_t k = l.dofbase[size_t(li)];
        for (int o = 0; o < off; ++o) k += mask_test(m, o);
        return k;
    };
    // an accessor read: the leaf's effective value, the default where no leaf exists
    auto coef_at = [&](int kind, const Coord& g) -> float {
        const int32_t li = l.lookup.find(leaf_origin_of(g));
        if (li < 0) return l.default_coef(kind);
        return l.coef[kind][size_t(li) * kLeafSize + size_t(voxel_offset(g.x & 7, g.y & 7, g.z & 7))];
    };
    auto put = [&](int64_t r, int64_t c, float v) {
        rows.push_back(int32_t(r));
        cols.push_back(int32_t(c));
        vals.push_back(v);
    };
    for (size_t i = 0; i < l.nleaf(); ++i) {
        int64_t k = l.dofbase[i];
        const Coord o = l.origin[i];
        for (int off = 0; off < kLeafSize; ++off) {
            if (!mask_test(&l.mask[i * 8], off)) continue;
            const Coord g{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
            put(k, k, l.coef[kCoefDiag][i * kLeafSize + size_t(off)]);
            for (int c = 0; c < 3; ++c) {
                Coord dn = g, up = g;
                dn[c] -= 1;
                up[c] += 1;
                const int64_t kd = dof_at(dn);
                if (kd >= 0) put(k, kd, coef_at(kCoefX + c, g));
                const int64_t ku = dof_at(up);
                if (ku >= 0) put(k, ku, coef_at(kCoefX + c, up));
            }
            ++k;
        }
    }
}

void build_poisson_rhs(const PoissonLevelHost& l0, const HostGrid& weights, const HostGrid& vx, const HostGrid& vy,
                       const HostGrid& vz, const HostGrid& solid_vel, float* out_rhs) {
    const HostGrid* vel[3] = {&vx, &vy, &vz};
    const float invdx = 1.0f / l0.dx;
    parallel_for(int64_t(l0.nleaf()), [&](int64_t b, int64_t e) {
        GridAccessor wa(weights), sa(solid_vel), va[3] = {GridAccessor(vx), GridAccessor(vy), GridAccessor(vz)};
        (void)vel;
        for (int64_t i = b; i < e; ++i) {
            const uint64_t* m = &l0.mask[i * 8];
            const Coord o = l0.origin[i];
            int64_t dof = l0.dofbase[i];
            for (int off = 0; off < kLeafSize; ++off) {
                if (!mask_test(m, off)) continue;
                const Coord g{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                float rhs = 0;
                bool nonzero_weight = false;
                for (int f = 0; f < 6; ++f) {
                    const int ch = f >> 1;
                    const bool positive = (f & 1) == 0;
                    Coord nc = g;
                    nc[ch] += positive ? 1 : 0;
                    const float weight = wa.value(nc, ch);
                    if (weight != 0.f) nonzero_weight = true;
                    const float v = va[ch].value(nc);
                    const float sv = sa.value(nc, ch);
                    const float flux = invdx * (weight * v + (1.0f - weight) * sv);
                    if (positive) rhs -= flux;
                    else rhs += flux;
                }
                if (!nonzero_weight) rhs = 0;
                out_rhs[dof++] = rhs;
            }
        }
    });
}

}  // namespace uaamg
