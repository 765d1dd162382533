#include "viscosity_host.h"

#include "build_timing.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <unordered_map>
#include <unordered_set>

#include "poisson_host.h"
#include "visc_volumes.h"
#include "volume_fraction.h"

namespace uaamg {

namespace {

// per-leaf loops: a leaf is 512 voxels of work, so even a few leaves are
// spread over the host threads
template <class Fn>
void leaf_for(int64_t n, Fn&& fn) {
    parallel_for(n, std::forward<Fn>(fn), 1);
}

inline Coord unit(int a) {
    Coord c{0, 0, 0};
    c[a] = 1;
    return c;
}
inline Coord shifted(Coord c, int a, int d) {
    c[a] += d;
    return c;
}
inline int off_of(const Coord& p) { return voxel_offset(p.x & 7, p.y & 7, p.z & 7); }

struct LeafSet {
    std::vector<Coord> org;
    LeafLookup lk;
    void finish() {
        std::sort(org.begin(), org.end(), tree_order_less);
        lk.build(org);
    }
    int32_t find(const Coord& o) const { return lk.find(o); }
    size_t size() const { return org.size(); }
};

// One-leaf-cached reads of per-leaf arrays (the role of OpenVDB's accessor cache).
template <class T, int N>
struct Acc {
    const LeafSet* ls;
    const T* data;
    T bg;
    uint64_t key = ~uint64_t(0);
    const T* p = nullptr;
    const T* leaf(const Coord& o) {
        const uint64_t k = pack_origin(o);
        if (k != key) {
            key = k;
            const int32_t i = ls->find(o);
            p = i < 0 ? nullptr : data + size_t(i) * N;
        }
        return p;
    }
};
struct FieldAcc : Acc<float, kLeafSize> {
    float get(const Coord& c) {
        const float* q = leaf(leaf_origin_of(c));
        return q ? q[off_of(c)] : bg;
    }
};
struct DofAcc : Acc<int32_t, kLeafSize> {
    int32_t get(const Coord& c) {
        const int32_t* q = leaf(leaf_origin_of(c));
        return q ? q[off_of(c)] : -1;
    }
};
struct GridAcc {
    const HostGrid* g;
    uint64_t key = ~uint64_t(0);
    const float* p = nullptr;
    float get(const Coord& c) {
        const Coord o = leaf_origin_of(c);
        const uint64_t k = pack_origin(o);
        if (k != key) {
            key = k;
            const int32_t i = g->lookup.find(o);
            p = i < 0 ? nullptr : g->leaf_values(size_t(i));
        }
        return p ? p[off_of(c) * g->ncomp] : g->background[0];
    }
};

// OpenVDB BoxSampler::sample restated: trilinear a + T((b - a) * w), z first
inline float lerp(float a, float b, double t) { return a + float(double(b - a) * t); }
float box_sample(GridAcc& acc, double px, double py, double pz) {
    if (acc.g->nleaf() == 0) return acc.g->background[0];
    const int ix = int(std::floor(px)), iy = int(std::floor(py)), iz = int(std::floor(pz));
    const double u = px - ix, v = py - iy, w = pz - iz;
    float d[2][2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) d[i][j][k] = acc.get({ix + i, iy + j, iz + k});
    return lerp(lerp(lerp(d[0][0][0], d[0][0][1], w), lerp(d[0][1][0], d[0][1][1], w), v),
                lerp(lerp(d[1][0][0], d[1][0][1], w), lerp(d[1][1][0], d[1][1][1], w), v), u);
}

// Leaves after OpenVDB dilateActiveValues of the active voxels (k iterations,
// 26 or 6 neighbours): the input leaves plus every neighbour leaf a dilated
// voxel lands in.
LeafSet dilate(const HostGrid& g, int k, bool nn26) {
    const size_t n = g.nleaf();
    std::vector<uint32_t> touched(n, 0);
    leaf_for(int64_t(n), [&](int64_t b, int64_t e) {
        for (int64_t li = b; li < e; ++li) {
            uint32_t t = 0;
            const uint64_t* m = &g.active[size_t(li) * 8];
            for (int off = 0; off < kLeafSize; ++off) {
                if (!mask_test(m, off)) continue;
                const int l[3] = {off >> 6, (off >> 3) & 7, off & 7};
                int dp[3], dn[3];
                for (int c = 0; c < 3; ++c) {
                    dp[c] = 8 - l[c];
                    dn[c] = l[c] + 1;
                }
                for (int d = 0; d < 27; ++d) {
                    if (d == 13 || ((t >> d) & 1u)) continue;
                    const int a[3] = {d / 9 - 1, (d / 3) % 3 - 1, d % 3 - 1};
                    int mx = 0, sm = 0;
                    for (int c = 0; c < 3; ++c) {
                        const int dd = a[c] > 0 ? dp[c] : (a[c] < 0 ? dn[c] : 0);
                        mx = std::max(mx, dd);
                        sm += dd;
                    }
                    if ((nn26 ? mx : sm) <= k) t |= 1u << d;
                }
            }
            touched[li] = t;
        }
    });
    LeafSet out;
    std::unordered_set<uint64_t> seen;
    seen.reserve(n * 4 + 16);
    for (size_t li = 0; li < n; ++li)
        if (seen.insert(pack_origin(g.origin[li])).second) out.org.push_back(g.origin[li]);
    for (size_t li = 0; li < n; ++li)
        for (int d = 0; d < 27; ++d)
            if ((touched[li] >> d) & 1u) {
                const Coord o{g.origin[li].x + 8 * (d / 9 - 1), g.origin[li].y + 8 * ((d / 3) % 3 - 1),
                              g.origin[li].z + 8 * (d % 3 - 1)};
                if (seen.insert(pack_origin(o)).second) out.org.push_back(o);
            }
    out.finish();
    return out;
}

// subcell fields: each liquid leaf (dilated) owns the 16^3 subcells of its volume
constexpr int kSub = 16, kSubSize = kSub * kSub * kSub;
struct SubField {
    const LeafSet* ls;
    std::vector<float> v;   // [n][16^3] x-major
    float bg;
    float get(const Coord& s) const {
        const Coord o = leaf_origin_of({s.x >> 1, s.y >> 1, s.z >> 1});
        const int32_t i = ls->find(o);
        if (i < 0) return bg;
        const int lx = s.x - 2 * o.x, ly = s.y - 2 * o.y, lz = s.z - 2 * o.z;
        return v[size_t(i) * kSubSize + (lx * kSub + ly) * kSub + lz];
    }
};

// coefficient bricks of a level under construction
struct CoefStore {
    OriginMap map[kVNumGrids];
    BrickVec pool;
    std::vector<uint64_t> on;
    int32_t nbricks = 0;
    // registers a brick (its storage comes with allocate())
    int32_t add(int gid, const Coord& o) {
        auto r = map[gid].emplace(pack_origin(o), nbricks);
        if (r.second) ++nbricks;
        return r.first;
    }
    // zeroed storage for every registered brick, after the last add()
    void allocate() {
        pool.resize(size_t(nbricks) * kLeafSize);
        on.assign(size_t(nbricks) * 8, 0);
        parallel_for(int64_t(nbricks), [&](int64_t b, int64_t e) {
            std::fill(pool.begin() + b * kLeafSize, pool.begin() + e * kLeafSize, 0.f);
        });
    }
    int32_t find(int gid, const Coord& o) const { return map[gid].find(pack_origin(o)); }
    void set(int32_t b, int off, float v) {
        pool[size_t(b) * kLeafSize + off] = v;
        on[size_t(b) * 8 + (off >> 6)] |= uint64_t(1) << (off & 63);
    }
    // value where the pattern leaves read 0 for anything not set (coarsening)
    float get0(int gid, const Coord& p) const {
        const int32_t b = find(gid, leaf_origin_of(p));
        return b < 0 ? 0.f : pool[size_t(b) * kLeafSize + off_of(p)];
    }
};

// get0 with a one-leaf cache per grid (the reads of one coarse voxel's
// children fall in one fine leaf)
struct StoreAcc {
    const CoefStore* cs;
    uint64_t key[kVNumGrids];
    const float* brick[kVNumGrids];
    explicit StoreAcc(const CoefStore& s) : cs(&s) {
        for (int g = 0; g < kVNumGrids; ++g) {
            key[g] = ~uint64_t(0);
            brick[g] = nullptr;
        }
    }
    float get0(int gid, const Coord& p) {
        const Coord o = leaf_origin_of(p);
        const uint64_t k = pack_origin(o);
        if (k != key[gid]) {
            key[gid] = k;
            const int32_t b = cs->map[gid].find(k);
            brick[gid] = b < 0 ? nullptr : &cs->pool[size_t(b) * kLeafSize];
        }
        return brick[gid] ? brick[gid][off_of(p)] : 0.f;
    }
};

// one level under construction
struct BuildLevel {
    float dx = 0, default_diag = 1, default_off = 1;
    float dflt[kVSlots] = {};
    int64_t ndof = 0, nchan[3] = {0, 0, 0};
    LeafSet dls[3];                 // DOF leaves per channel (each holds >= 1 DOF)
    std::vector<int32_t> didx[3];   // [n][512] DOF index or -1
    CoefStore cs;
    DofAcc dof_acc(int c) const {
        DofAcc a;
        a.ls = &dls[c];
        a.data = didx[c].data();
        a.bg = -1;
        return a;
    }
};

void set_defaults(BuildLevel& L) {
    const float d = L.default_diag, o = L.default_off;
    for (int c = 0; c < 3; ++c) {
        L.dflt[kVDiag + c] = d;
        L.dflt[kVInvDiag + c] = 1.0f / d;
        for (int a = 0; a < 3; ++a) L.dflt[vneg(c, a)] = a == c ? -2.0f * o : -o;
        L.dflt[kVSpai0 + c] = 0.f;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) L.dflt[vcross(i, j, k)] = (j + k) % 2 ? -o : o;
}

// keep only leaves with a DOF, numbered in tree order from `ndof` on
void number_channel(const LeafSet& cand, std::vector<uint64_t>& on, BuildLevel& L, int c) {
    std::vector<int64_t> cnt(cand.size());
    for (size_t i = 0; i < cand.size(); ++i) cnt[i] = mask_count(&on[i * 8]);
    L.dls[c] = LeafSet();
    for (size_t i = 0; i < cand.size(); ++i)
        if (cnt[i]) L.dls[c].org.push_back(cand.org[i]);
    L.dls[c].lk.build(L.dls[c].org);   // cand is tree ordered, so is the subset
    L.didx[c].assign(L.dls[c].size() * kLeafSize, -1);
    size_t k = 0;
    for (size_t i = 0; i < cand.size(); ++i) {
        if (!cnt[i]) continue;
        int32_t* d = &L.didx[c][k * kLeafSize];
        for (int off = 0; off < kLeafSize; ++off)
            if (mask_test(&on[i * 8], off)) d[off] = int32_t(L.ndof++);
        ++k;
    }
}

// leaves at offsets {-1,0}^3 of the DOF leaves of two channels (cross-term positions)
LeafSet cross_candidates(const BuildLevel& L, int c1, int c2) {
    LeafSet s;
    std::unordered_set<uint64_t> seen;
    for (int ch : {c1, c2})
        for (const Coord& o : L.dls[ch].org)
            for (int d = 0; d < 8; ++d) {
                const Coord q{o.x - 8 * ((d >> 2) & 1), o.y - 8 * ((d >> 1) & 1), o.z - 8 * (d & 1)};
                if (seen.insert(pack_origin(q)).second) s.org.push_back(q);
            }
    s.finish();
    return s;
}

// --------------------------------------------------------------- finest
bool build_finest(const HostGrid& liquid, const HostGrid& solid, const HostGrid& mu, float dt, float rho, float dx,
                  int volume_device, BuildLevel& L, ViscGeometry& G, std::string& err) {
    L.dx = dx;
    G.dx = dx;
    G.dt = dt;
    G.rho = rho;
    StageClock clock;
    G.mu = mu;
    G.mu.lookup.build(G.mu.origin);

    const LeafSet sub_ls = dilate(liquid, 1, true);
    const LeafSet vol_ls = dilate(liquid, 2, true);
    G.vol_origin = vol_ls.org;
    G.vol_lookup.build(G.vol_origin);
    const size_t nv = vol_ls.size();
    if (volume_device >= 0) {
        // the subcell SDF, subcell volumes and voxel / face / edge volumes on the device
        if (visc_volumes_device(volume_device, liquid, sub_ls.org, vol_ls.org, dx, G.voxel_vol, G.face_vol,
                                G.edge_vol, err) != 0)
            return false;
    } else {
    // subcell SDF: corner s of the subcell lattice sits at liquid index (s-1)/2
    SubField sdf{&sub_ls, std::vector<float>(sub_ls.size() * kSubSize), 1.5f * dx};
    leaf_for(int64_t(sub_ls.size()), [&](int64_t b, int64_t e) {
        GridAcc acc{&liquid};
        float blk[10][10][10];
        for (int64_t li = b; li < e; ++li) {
            const Coord o = sub_ls.org[li];
            for (int i = 0; i < 10; ++i)
                for (int j = 0; j < 10; ++j)
                    for (int k = 0; k < 10; ++k) blk[i][j][k] = acc.get({o.x - 1 + i, o.y - 1 + j, o.z - 1 + k});
            float* out = &sdf.v[size_t(li) * kSubSize];
            for (int lx = 0; lx < kSub; ++lx)
                for (int ly = 0; ly < kSub; ++ly)
                    for (int lz = 0; lz < kSub; ++lz) {
                        // s = 2o + l; floor((s-1)/2) - (o-1) = (l+1)>>1; weight 0 (odd s) or 1/2
                        const int ix = (lx + 1) >> 1, iy = (ly + 1) >> 1, iz = (lz + 1) >> 1;
                        const double u = (lx & 1) ? 0.0 : 0.5, v = (ly & 1) ? 0.0 : 0.5, w = (lz & 1) ? 0.0 : 0.5;
                        const float smp =
                            lerp(lerp(lerp(blk[ix][iy][iz], blk[ix][iy][iz + 1], w),
                                      lerp(blk[ix][iy + 1][iz], blk[ix][iy + 1][iz + 1], w), v),
                                 lerp(lerp(blk[ix + 1][iy][iz], blk[ix + 1][iy][iz + 1], w),
                                      lerp(blk[ix + 1][iy + 1][iz], blk[ix + 1][iy + 1][iz + 1], w), v),
                                 u);
                        out[(lx * kSub + ly) * kSub + lz] = smp < sdf.bg ? smp : sdf.bg;
                    }
        }
    });
    // subcell liquid volumes from the 8 corners
    SubField svol{&sub_ls, std::vector<float>(sub_ls.size() * kSubSize, 0.f), 0.f};
    leaf_for(int64_t(sub_ls.size()), [&](int64_t b, int64_t e) {
        std::vector<float> c17(17 * 17 * 17);
        for (int64_t li = b; li < e; ++li) {
            const Coord o = sub_ls.org[li];
            const float* own = &sdf.v[size_t(li) * kSubSize];
            for (int i = 0; i < 17; ++i)
                for (int j = 0; j < 17; ++j)
                    for (int k = 0; k < 17; ++k)
                        c17[(i * 17 + j) * 17 + k] = (i < 16 && j < 16 && k < 16)
                                                         ? own[(i * kSub + j) * kSub + k]
                                                         : sdf.get({2 * o.x + i, 2 * o.y + j, 2 * o.z + k});
            float* out = &svol.v[size_t(li) * kSubSize];
            for (int i = 0; i < 16; ++i)
                for (int j = 0; j < 16; ++j)
                    for (int k = 0; k < 16; ++k) {
                        auto P = [&](int a, int bb, int c) { return c17[((i + a) * 17 + j + bb) * 17 + k + c]; };
                        const float f = visc_volume_fraction(P(0, 0, 0), P(1, 0, 0), P(0, 1, 0), P(1, 1, 0),
                                                             P(0, 0, 1), P(1, 0, 1), P(0, 1, 1), P(1, 1, 1));
                        out[(i * kSub + j) * kSub + k] = f > 0 ? f : 0.f;
                    }
        }
    });

    // voxel / face / edge volumes on the liquid leaves dilated twice
    G.voxel_vol.assign(nv * kLeafSize, 0.f);
    for (int c = 0; c < 3; ++c) {
        G.face_vol[c].assign(nv * kLeafSize, 0.f);
        G.edge_vol[c].assign(nv * kLeafSize, 0.f);
    }
    leaf_for(int64_t(nv), [&](int64_t b, int64_t e) {
        std::vector<float> s18(18 * 18 * 18);
        for (int64_t li = b; li < e; ++li) {
            const Coord o = vol_ls.org[li];
            // subcells 2o-1 .. 2o+16
            for (int i = 0; i < 18; ++i)
                for (int j = 0; j < 18; ++j)
                    for (int k = 0; k < 18; ++k)
                        s18[(i * 18 + j) * 18 + k] = svol.get({2 * o.x - 1 + i, 2 * o.y - 1 + j, 2 * o.z - 1 + k});
            auto S = [&](int x, int y, int z) { return s18[((x + 1) * 18 + y + 1) * 18 + z + 1]; };
            for (int off = 0; off < kLeafSize; ++off) {
                const int vx = off >> 6, vy = (off >> 3) & 7, vz = off & 7;
                float vc = 0, vf[3] = {0, 0, 0}, ve[3] = {0, 0, 0};
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j)
                        for (int k = 0; k < 2; ++k) {
                            const int a[3] = {2 * vx + i, 2 * vy + j, 2 * vz + k};
                            vc += S(a[0], a[1], a[2]);
                            for (int c = 0; c < 3; ++c) {
                                int f[3] = {a[0], a[1], a[2]};
                                f[c] -= 1;
                                vf[c] += S(f[0], f[1], f[2]);
                                int g[3] = {a[0] - 1, a[1] - 1, a[2] - 1};
                                g[c] += 1;
                                ve[c] += S(g[0], g[1], g[2]);
                            }
                        }
                const size_t k = size_t(li) * kLeafSize + off;
                G.voxel_vol[k] = vc > 0 ? vc * 0.125f : 0.f;
                for (int c = 0; c < 3; ++c) {
                    G.face_vol[c][k] = vf[c] > 0 ? vf[c] * 0.125f : 0.f;
                    G.edge_vol[c][k] = ve[c] > 0 ? ve[c] * 0.125f : 0.f;
                }
            }
        }
    });
    }
    clock.lap("viscosity level 0: volume stage");
    LeafSet vls;
    vls.org = G.vol_origin;
    vls.lk.build(vls.org);
    auto vacc = [&](const std::vector<float>& v) {
        FieldAcc a;
        a.ls = &vls;
        a.data = v.data();
        a.bg = 0.f;
        return a;
    };

    // solid faces and DOF marking on the liquid leaves dilated once (6 neighbours)
    const LeafSet dof_ls = dilate(liquid, 1, false);
    G.dof_origin = dof_ls.org;
    G.dof_lookup.build(G.dof_origin);
    const size_t nd = dof_ls.size();
    std::vector<uint64_t> dof_on[3];
    for (int i = 0; i < 3; ++i) {
        G.solid_on[i].assign(nd * 8, 0);
        dof_on[i].assign(nd * 8, 0);
    }
    leaf_for(int64_t(nd), [&](int64_t b, int64_t e) {
        GridAcc sacc{&solid};
        FieldAcc va = vacc(G.voxel_vol);
        FieldAcc fa[3] = {vacc(G.face_vol[0]), vacc(G.face_vol[1]), vacc(G.face_vol[2])};
        FieldAcc ea[3] = {vacc(G.edge_vol[0]), vacc(G.edge_vol[1]), vacc(G.edge_vol[2])};
        for (int64_t li = b; li < e; ++li) {
            const Coord o = dof_ls.org[li];
            for (int i = 0; i < 3; ++i) {
                const int d1 = (i + 1) % 3, d2 = (i + 2) % 3;
                for (int off = 0; off < kLeafSize; ++off) {
                    const Coord v{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    double p[3] = {v.x + 0.5, v.y + 0.5, v.z + 0.5};
                    p[i] = v[i];
                    if (box_sample(sacc, p[0], p[1], p[2]) < 0) {
                        mask_set(&G.solid_on[i][size_t(li) * 8], off);
                        continue;
                    }
                    bool on;
                    if (fa[i].get(v) > 0) on = true;
                    else if (va.get(v) > 0 || va.get(shifted(v, i, -1)) > 0) on = true;
                    else
                        on = ea[d1].get(v) > 0 || ea[d1].get(v + unit(d2)) > 0 || ea[d2].get(v) > 0 ||
                             ea[d2].get(v + unit(d1)) > 0;
                    if (on) mask_set(&dof_on[i][size_t(li) * 8], off);
                }
            }
        }
    });
    for (int i = 0; i < 3; ++i) {
        const int64_t before = L.ndof;
        number_channel(dof_ls, dof_on[i], L, i);
        L.nchan[i] = L.ndof - before;
        // geometry keeps the index on the unpruned leaves for the right-hand side
        G.dof_index[i].assign(nd * kLeafSize, -1);
        for (size_t k = 0, li = 0; li < nd; ++li) {
            if (!mask_count(&dof_on[i][li * 8])) continue;
            std::copy(&L.didx[i][k * kLeafSize], &L.didx[i][(k + 1) * kLeafSize], &G.dof_index[i][li * kLeafSize]);
            ++k;
        }
    }

    clock.lap("viscosity level 0: DOF marking + numbering");
    // coefficients (finest_level_matrix_setter, :539-908)
    const float factor = dt / (rho * dx * dx);
    L.default_off = factor * mu.background[0];
    L.default_diag = 1.0f + 8.0f * L.default_off;
    set_defaults(L);
    const float eps = 1e-1f;
    // register bricks: channel bricks on the channel's DOF leaves, cross
    // bricks on the candidate leaves of their two channels
    for (int c = 0; c < 3; ++c)
        for (const Coord& o : L.dls[c].org) {
            L.cs.add(kVDiag + c, o);
            L.cs.add(kVInvDiag + c, o);
            for (int a = 0; a < 3; ++a) L.cs.add(vneg(c, a), o);
        }
    LeafSet cand[3];
    for (int d0 = 0; d0 < 3; ++d0) {
        cand[d0] = cross_candidates(L, (d0 + 1) % 3, (d0 + 2) % 3);
        for (const Coord& o : cand[d0].org)
            for (int t = 0; t < 4; ++t) L.cs.add(vcross(d0, t >> 1, t & 1), o);
    }
    L.cs.allocate();
    for (int c = 0; c < 3; ++c) {
        const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const LeafSet& ls = L.dls[c];
        leaf_for(int64_t(ls.size()), [&](int64_t b, int64_t e) {
            FieldAcc va = vacc(G.voxel_vol);
            FieldAcc fa = vacc(G.face_vol[d0]);
            FieldAcc e1 = vacc(G.edge_vol[d1]), e2 = vacc(G.edge_vol[d2]);
            DofAcc da = L.dof_acc(d0);
            GridAcc ma{&G.mu};
            for (int64_t li = b; li < e; ++li) {
                const Coord o = ls.org[li];
                const int32_t bd = L.cs.find(kVDiag + c, o), bi = L.cs.find(kVInvDiag + c, o);
                const int32_t bn[3] = {L.cs.find(vneg(c, d0), o), L.cs.find(vneg(c, d1), o), L.cs.find(vneg(c, d2), o)};
                const int32_t* idx = &L.didx[c][size_t(li) * kLeafSize];
                for (int off = 0; off < kLeafSize; ++off) {
                    if (idx[off] < 0) continue;
                    const Coord v{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    float diag = 0;
                    float fv = fa.get(v);
                    if (fv < eps) fv = eps;
                    diag += fv;
                    {
                        const float t = 2 * factor * va.get(v) * ma.get(v);
                        diag += t;
                    }
                    {
                        const Coord n = shifted(v, d0, -1);
                        const float t = 2 * factor * va.get(n) * ma.get(n);
                        diag += t;
                        if (da.get(n) != -1) L.cs.set(bn[0], off, -t);
                    }
                    float fc[3] = {float(v.x), float(v.y), float(v.z)};
                    fc[d0] -= 0.5f;
                    auto mus = [&](int a, float da_) {
                        float p[3] = {fc[0], fc[1], fc[2]};
                        p[a] += da_;
                        return box_sample(ma, p[0], p[1], p[2]);
                    };
                    {
                        const float t = factor * e2.get(v + unit(d1)) * mus(d1, 0.5f);
                        diag += t;
                    }
                    {
                        const float t = factor * e2.get(v) * mus(d1, -0.5f);
                        diag += t;
                        if (da.get(shifted(v, d1, -1)) != -1) L.cs.set(bn[1], off, -t);
                    }
                    {
                        const float t = factor * e1.get(v + unit(d2)) * mus(d2, 0.5f);
                        diag += t;
                    }
                    {
                        const float t = factor * e1.get(v) * mus(d2, -0.5f);
                        diag += t;
                        if (da.get(shifted(v, d2, -1)) != -1) L.cs.set(bn[2], off, -t);
                    }
                    L.cs.set(bd, off, diag);
                    L.cs.set(bi, off, diag != 0 ? 1.0f / diag : 0.f);
                }
            }
        });
    }
    for (int d0 = 0; d0 < 3; ++d0) {
        const int d1 = (d0 + 1) % 3, d2 = (d0 + 2) % 3;
        const LeafSet& ls = cand[d0];
        leaf_for(int64_t(ls.size()), [&](int64_t b, int64_t e) {
            FieldAcc ea = vacc(G.edge_vol[d0]);
            DofAcc a1 = L.dof_acc(d1), a2 = L.dof_acc(d2);
            GridAcc ma{&G.mu};
            for (int64_t li = b; li < e; ++li) {
                const Coord o = ls.org[li];
                int32_t bx[4];
                for (int t = 0; t < 4; ++t) bx[t] = L.cs.find(vcross(d0, t >> 1, t & 1), o);
                for (int off = 0; off < kLeafSize; ++off) {
                    const Coord v{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    const bool n1 = a1.get(v) != -1, p1 = a1.get(v + unit(d1)) != -1;
                    const bool n2 = a2.get(v) != -1, p2 = a2.get(v + unit(d2)) != -1;
                    const bool pr[4] = {n1 && n2, n1 && p2, p1 && n2, p1 && p2};
                    for (int t = 0; t < 4; ++t) {
                        if (!pr[t]) continue;
                        const int j = t >> 1, k = t & 1;
                        Coord eo = v;
                        if (j) eo = eo + unit(d1);
                        if (k) eo = eo + unit(d2);
                        float p[3] = {float(v.x), float(v.y), float(v.z)};
                        p[d1] += j ? 0.5f : -0.5f;
                        p[d2] += k ? 0.5f : -0.5f;
                        const float f = (j + k) % 2 ? -factor : factor;
                        L.cs.set(bx[t], off, f * ea.get(eo) * box_sample(ma, p[0], p[1], p[2]));
                    }
                }
            }
        });
    }
    return true;
}

// ------------------------------------------------------------- coarse
void build_coarse(const BuildLevel& F, BuildLevel& C) {
    StageClock clock;
    C.dx = F.dx * 2.0f;
    C.default_diag = 1.0f;
    C.default_off = 1.0f;
    set_defaults(C);
    for (int i = 0; i < 3; ++i) {
        LeafSet cl;
        std::unordered_set<uint64_t> seen;
        for (const Coord& o : F.dls[i].org) {
            const Coord co = leaf_origin_of({o.x / 2, o.y / 2, o.z / 2});
            if (seen.insert(pack_origin(co)).second) cl.org.push_back(co);
        }
        cl.finish();
        std::vector<uint64_t> on(cl.size() * 8, 0);
        leaf_for(int64_t(cl.size()), [&](int64_t b, int64_t e) {
            DofAcc fa = F.dof_acc(i);
            for (int64_t li = b; li < e; ++li) {
                const Coord o = cl.org[li];
                for (int off = 0; off < kLeafSize; ++off) {
                    const Coord c{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    bool any = false;
                    for (int a = 0; a < 2 && !any; ++a)
                        for (int bb = 0; bb < 2 && !any; ++bb)
                            for (int d = 0; d < 2 && !any; ++d)
                                any = fa.get({2 * c.x + a, 2 * c.y + bb, 2 * c.z + d}) != -1;
                    if (any) mask_set(&on[size_t(li) * 8], off);
                }
            }
        });
        const int64_t before = C.ndof;
        number_channel(cl, on, C, i);
        C.nchan[i] = C.ndof - before;
    }
    clock.lap("viscosity coarse: topology");
    for (int c = 0; c < 3; ++c)
        for (const Coord& o : C.dls[c].org) {
            C.cs.add(kVDiag + c, o);
            C.cs.add(kVInvDiag + c, o);
            for (int a = 0; a < 3; ++a) C.cs.add(vneg(c, a), o);
        }
    LeafSet cand[3];
    for (int d0 = 0; d0 < 3; ++d0) {
        cand[d0] = cross_candidates(C, (d0 + 1) % 3, (d0 + 2) % 3);
        for (const Coord& o : cand[d0].org)
            for (int t = 0; t < 4; ++t) C.cs.add(vcross(d0, t >> 1, t & 1), o);
    }
    C.cs.allocate();
    clock.lap("viscosity coarse: brick store");
    const float rap = 0.125f;
    for (int i = 0; i < 3; ++i) {
        const int d0 = i, d1 = (i + 1) % 3, d2 = (i + 2) % 3;
        const int dd[3] = {d0, d1, d2};
        const LeafSet& ls = C.dls[i];
        leaf_for(int64_t(ls.size()), [&](int64_t b, int64_t e) {
            DofAcc fa = F.dof_acc(i), ca = C.dof_acc(i);
            DofAcc fn = F.dof_acc(i);
            StoreAcc fs(F.cs), fsn(F.cs);
            for (int64_t li = b; li < e; ++li) {
                const Coord o = ls.org[li];
                const int32_t bd = C.cs.find(kVDiag + i, o), bi = C.cs.find(kVInvDiag + i, o);
                int32_t bn[3];
                for (int k = 0; k < 3; ++k) bn[k] = C.cs.find(vneg(i, dd[k]), o);
                const int32_t* idx = &C.didx[i][size_t(li) * kLeafSize];
                for (int off = 0; off < kLeafSize; ++off) {
                    if (idx[off] < 0) continue;
                    const Coord c{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    float nd = 0.f, nn[3] = {0.f, 0.f, 0.f};
                    for (int ii = 0; ii < 2; ++ii)
                        for (int jj = 0; jj < 2; ++jj)
                            for (int kk = 0; kk < 2; ++kk) {
                                Coord f{2 * c.x, 2 * c.y, 2 * c.z};
                                f[d0] += ii;
                                f[d1] += jj;
                                f[d2] += kk;
                                if (fa.get(f) == -1) continue;
                                nd += fs.get0(kVDiag + i, f);
                                const int loc[3] = {ii, jj, kk};
                                for (int k = 0; k < 3; ++k) {
                                    const float t = fs.get0(vneg(i, dd[k]), f);
                                    if (fn.get(shifted(f, dd[k], -1)) == -1) continue;
                                    if (loc[k] == 0) nn[k] += t;
                                    else nd += t * 2.0f;
                                }
                            }
                    C.cs.set(bd, off, rap * nd);
                    C.cs.set(bi, off, nd != 0 ? 1.0f / (rap * nd) : 0.f);
                    for (int k = 0; k < 3; ++k)
                        if (ca.get(shifted(c, dd[k], -1)) != -1) C.cs.set(bn[k], off, rap * nn[k]);
                }
            }
        });
        const LeafSet& cl = cand[i];
        const int fx[4] = {vcross(i, 0, 0), vcross(i, 0, 1), vcross(i, 1, 0), vcross(i, 1, 1)};
        leaf_for(int64_t(cl.size()), [&](int64_t b, int64_t e) {
            DofAcc a1 = C.dof_acc(d1), a2 = C.dof_acc(d2);
            StoreAcc fs(F.cs);
            for (int64_t li = b; li < e; ++li) {
                const Coord o = cl.org[li];
                int32_t bx[4];
                for (int t = 0; t < 4; ++t) bx[t] = C.cs.find(fx[t], o);
                for (int off = 0; off < kLeafSize; ++off) {
                    const Coord c{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    const bool n1 = a1.get(c) != -1, p1 = a1.get(c + unit(d1)) != -1;
                    const bool n2 = a2.get(c) != -1, p2 = a2.get(c + unit(d2)) != -1;
                    const bool pr[4] = {n1 && n2, n1 && p2, p1 && n2, p1 && p2};
                    for (int t = 0; t < 4; ++t) {
                        if (!pr[t]) continue;
                        float s = 0.f;
                        for (int ii = 0; ii < 2; ++ii)
                            for (int jj = 0; jj < 2; ++jj)
                                for (int kk = 0; kk < 2; ++kk) {
                                    Coord f{2 * c.x, 2 * c.y, 2 * c.z};
                                    f[d0] += ii;
                                    f[d1] += jj;
                                    f[d2] += kk;
                                    switch (t) {
                                        case 0:
                                            s += fs.get0(fx[0], f);
                                            if (kk == 0) s += fs.get0(fx[1], f);
                                            if (jj == 0) s += fs.get0(fx[2], f);
                                            if (kk == 0 && jj == 0) s += fs.get0(fx[3], f);
                                            break;
                                        case 1:
                                            if (kk == 1) s += fs.get0(fx[1], f);
                                            if (jj == 0 && kk == 1) s += fs.get0(fx[3], f);
                                            break;
                                        case 2:
                                            if (jj == 1) s += fs.get0(fx[2], f);
                                            if (jj == 1 && kk == 0) s += fs.get0(fx[3], f);
                                            break;
                                        default:
                                            if (jj == 1 && kk == 1) s += fs.get0(fx[3], f);
                                    }
                                }
                        C.cs.set(bx[t], off, rap * s);
                    }
                }
            }
        });
    }
}

// trim_default_nodes (:1700-1722): bricks whose active values all lie within
// |eps| of the default read as the default
void trim(BuildLevel& L) {
    const float eps = 1e-3f;
    for (int gid = 0; gid < kVSpai0; ++gid) {
        float dv = L.dflt[gid], ep;
        if (gid >= kVDiag && gid < kVInvDiag) ep = L.default_diag * eps;
        else if (gid < kVNeg) ep = 1.0f / L.default_diag * eps;
        else ep = dv * eps;
        ep = std::abs(ep);
        std::vector<std::pair<uint64_t, int32_t>> bricks;
        L.cs.map[gid].for_each([&](uint64_t key, int32_t b) { bricks.emplace_back(key, b); });
        std::vector<uint8_t> drop(bricks.size(), 0);
        leaf_for(int64_t(bricks.size()), [&](int64_t lo, int64_t hi) {
            for (int64_t t = lo; t < hi; ++t) {
                const int32_t b = bricks[size_t(t)].second;
                const float* v = &L.cs.pool[size_t(b) * kLeafSize];
                const uint64_t* on = &L.cs.on[size_t(b) * 8];
                float me = 0;
                for (int off = 0; off < kLeafSize; ++off)
                    if (mask_test(on, off)) me = std::max(me, std::abs(v[off] - dv));
                drop[size_t(t)] = me <= ep;
            }
        });
        for (size_t t = 0; t < bricks.size(); ++t)
            if (drop[t]) L.cs.map[gid].remove(bricks[t].first);
    }
}

// drop bricks with no active value (coarse levels keep zeros there, which the
// operator multiplies with zero velocities only)
void drop_empty(BuildLevel& L) {
    for (int gid = 0; gid < kVNumGrids; ++gid) {
        std::vector<uint64_t> gone;
        L.cs.map[gid].for_each([&](uint64_t key, int32_t b) {
            if (!mask_count(&L.cs.on[size_t(b) * 8])) gone.push_back(key);
        });
        for (uint64_t key : gone) L.cs.map[gid].remove(key);
    }
}

void finalize(const BuildLevel& B, int level, float dt, float rho, ViscLevelHost& L) {
    L.level = level;
    L.dx = B.dx;
    L.dt = dt;
    L.rho = rho;
    L.default_diag = B.default_diag;
    L.default_off = B.default_off;
    std::copy(B.dflt, B.dflt + kVSlots, L.dflt);
    L.ndof = B.ndof;
    for (int c = 0; c < 3; ++c) L.nchan[c] = B.nchan[c];
    std::unordered_set<uint64_t> seen;
    std::vector<Coord> dl, extra;
    for (int c = 0; c < 3; ++c)
        for (const Coord& o : B.dls[c].org)
            if (seen.insert(pack_origin(o)).second) dl.push_back(o);
    std::sort(dl.begin(), dl.end(), tree_order_less);
    std::unordered_map<uint64_t, Coord> all;
    for (int gid = 0; gid < kVNumGrids; ++gid)
        B.cs.map[gid].for_each([&](uint64_t key, int32_t) {
            if (!seen.count(key)) {
                const int64_t m = (int64_t(1) << 21);
                auto unpack = [&](uint64_t v) {
                    int64_t t = int64_t(v & uint64_t(m - 1));
                    if (t >= m / 2) t -= m;
                    return int32_t(t * 8);
                };
                all.emplace(key, Coord{unpack(key >> 42), unpack(key >> 21), unpack(key)});
            }
        });
    for (const auto& kv : all) extra.push_back(kv.second);
    std::sort(extra.begin(), extra.end(), tree_order_less);
    L.origin = dl;
    L.nleaf_dof = dl.size();
    L.origin.insert(L.origin.end(), extra.begin(), extra.end());
    L.lookup.build(L.origin);
    const size_t n = L.origin.size(), nd = L.nleaf_dof;
    L.mask.assign(nd * 3 * 8, 0);
    L.dofbase.assign(nd * 3, -1);
    for (int c = 0; c < 3; ++c) L.chan_leaves[c].clear();
    for (size_t i = 0; i < nd; ++i)
        for (int c = 0; c < 3; ++c) {
            const int32_t k = B.dls[c].find(L.origin[i]);
            if (k < 0) continue;
            const int32_t* idx = &B.didx[c][size_t(k) * kLeafSize];
            for (int off = 0; off < kLeafSize; ++off)
                if (idx[off] >= 0) {
                    if (L.dofbase[i * 3 + c] < 0) L.dofbase[i * 3 + c] = idx[off];
                    mask_set(&L.mask[(i * 3 + c) * 8], off);
                }
            L.chan_leaves[c].push_back(int32_t(i));
        }
    L.nb27.assign(n * 27, -1);
    leaf_for(int64_t(n), [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i)
            for (int d = 0; d < 27; ++d)
                L.nb27[size_t(i) * 27 + d] = L.lookup.find({L.origin[i].x + 8 * (d / 9 - 1),
                                                            L.origin[i].y + 8 * ((d / 3) % 3 - 1),
                                                            L.origin[i].z + 8 * (d % 3 - 1)});
    });
    L.slot.assign(n * kVSlots, -1);
    // slots: the bricks of each leaf in grid order, leaves in level order
    std::vector<int32_t> src;
    for (size_t i = 0; i < n; ++i)
        for (int gid = 0; gid < kVNumGrids; ++gid) {
            const int32_t b = B.cs.find(gid, L.origin[i]);
            if (b < 0) continue;
            L.slot[i * kVSlots + gid] = int32_t(src.size());
            src.push_back(b);
        }
    L.pool.resize(src.size() * kLeafSize);
    L.pool_on.resize(src.size() * 8);
    parallel_for(int64_t(src.size()), [&](int64_t lo, int64_t hi) {
        for (int64_t t = lo; t < hi; ++t) {
            const size_t b = size_t(src[size_t(t)]);
            std::copy(&B.cs.pool[b * kLeafSize], &B.cs.pool[(b + 1) * kLeafSize], &L.pool[size_t(t) * kLeafSize]);
            std::copy(&B.cs.on[b * 8], &B.cs.on[(b + 1) * 8], &L.pool_on[size_t(t) * 8]);
        }
    });
}

void link_levels(ViscLevelHost& F, ViscLevelHost& C) {
    F.parent.assign(F.nleaf_dof, -1);
    C.child.assign(C.nleaf_dof * 8, -1);
    for (size_t i = 0; i < F.nleaf_dof; ++i) {
        const Coord o = F.origin[i];
        const Coord h{o.x / 2, o.y / 2, o.z / 2};
        const int32_t j = C.lookup.find(leaf_origin_of(h));
        if (j < 0 || size_t(j) >= C.nleaf_dof) continue;
        const int oct = (((h.x & 4) >> 2) << 2) | (((h.y & 4) >> 2) << 1) | ((h.z & 4) >> 2);
        F.parent[i] = j * 8 + oct;
        C.child[size_t(j) * 8 + oct] = int32_t(i);
    }
}

}  // namespace

float visc_volume_fraction(float p000, float p100, float p010, float p110, float p001, float p101, float p011,
                           float p111) {
    return volume_fraction8(p000, p100, p010, p110, p001, p101, p011, p111);
}

bool build_visc_hierarchy(const HostGrid& liquid, const HostGrid& solid, const HostGrid& mu, float dt, float rho,
                          float dx, int64_t max_coarsest_dof, bool do_trim, ViscHierarchyHost& out, int volume_device,
                          std::string* err) {
    std::string e;
    if (!err) err = &e;
    if (liquid.ncomp != 1 || solid.ncomp != 1 || mu.ncomp != 1) {
        *err = "grids must have one component";
        return false;
    }
    StageClock clock;
    out = ViscHierarchyHost();
    std::vector<BuildLevel> B(1);
    if (!build_finest(liquid, solid, mu, dt, rho, dx, volume_device, B[0], out.geo, *err)) {
        if (err->empty()) *err = "level-0 construction failed";
        return false;
    }
    clock.lap("viscosity level 0");
    while (B.back().ndof > max_coarsest_dof) {
        B.emplace_back();
        build_coarse(B[B.size() - 2], B.back());
        if (B.size() > 40) return false;
        clock.lap("viscosity coarse level");
    }
    if (do_trim) trim(B[0]);
    else drop_empty(B[0]);
    for (size_t l = 1; l < B.size(); ++l) drop_empty(B[l]);
    clock.lap("viscosity trim");
    out.levels.resize(B.size());
    for (size_t l = 0; l < B.size(); ++l) finalize(B[l], int(l), dt, rho, out.levels[l]);
    if (std::getenv("UAAMG_BUILD_TIMING"))
        for (size_t l = 0; l < B.size(); ++l)
            std::fprintf(stderr, "[uaamg build] viscosity level %zu: %zu leaves (%zu with DOFs), %zu bricks\n", l,
                         out.levels[l].nleaf(), out.levels[l].nleaf_dof, out.levels[l].pool.size() / kLeafSize);
    for (size_t l = 0; l + 1 < B.size(); ++l) link_levels(out.levels[l], out.levels[l + 1]);
    clock.lap("viscosity finalize + links");
    visc_explicit_matrix(out.levels.back(), out.cptr, out.ccol, out.cval);
    clock.lap("viscosity coarsest matrix");
    return true;
}

// ---------------------------------------------------- SPAI0 / explicit
namespace {
// coefficient the level operator reads at a global voxel
inline float coef_at(const ViscLevelHost& L, int gid, const Coord& p) {
    const int32_t i = L.lookup.find(leaf_origin_of(p));
    if (i < 0) return L.dflt[gid];
    return L.coef(gid, size_t(i), off_of(p));
}
inline int32_t dof_at(const ViscLevelHost& L, int c, const Coord& p) {
    const int32_t i = L.lookup.find(leaf_origin_of(p));
    if (i < 0 || size_t(i) >= L.nleaf_dof) return -1;
    const uint64_t* m = &L.mask[(size_t(i) * 3 + c) * 8];
    const int off = off_of(p);
    if (!mask_test(m, off)) return -1;
    int before = 0;
    for (int w = 0; w < (off >> 6); ++w) before += __builtin_popcountll(m[w]);
    before += __builtin_popcountll(m[off >> 6] & ((uint64_t(1) << (off & 63)) - 1));
    return int32_t(L.dofbase[size_t(i) * 3 + c] + before);
}
}  // namespace

void build_visc_spai0(ViscLevelHost& L) {
    if (L.spai0) return;
    std::vector<std::vector<float>> vals(3, std::vector<float>(L.nleaf_dof * kLeafSize, 0.f));
    for (int c = 0; c < 3; ++c) {
        const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        leaf_for(int64_t(L.chan_leaves[c].size()), [&](int64_t b, int64_t e) {
            for (int64_t t = b; t < e; ++t) {
                const size_t i = size_t(L.chan_leaves[c][t]);
                const Coord o = L.origin[i];
                for (int off = 0; off < kLeafSize; ++off) {
                    if (!L.is_dof(i, c, off)) continue;
                    const Coord v{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    float den = 0.f, num;
                    const float dg = L.coef(kVDiag + c, i, off);
                    den += dg * dg;
                    num = dg;
                    for (int a = 0; a < 3; ++a) {
                        const Coord up = shifted(v, a, 1), dn = shifted(v, a, -1);
                        if (dof_at(L, d0, up) >= 0) {
                            const float q = coef_at(L, vneg(c, a), up);
                            den += q * q;
                        }
                        if (dof_at(L, d0, dn) >= 0) {
                            const float q = L.coef(vneg(c, a), i, off);
                            den += q * q;
                        }
                    }
                    auto sq = [&](bool on, int gid, const Coord& p) {
                        if (on) {
                            const float q = coef_at(L, gid, p);
                            den += q * q;
                        }
                    };
                    const Coord m0 = shifted(v, d0, -1);
                    sq(dof_at(L, d1, v) >= 0, vcross(d2, 0, 0), v);
                    sq(dof_at(L, d1, v + unit(d1)) >= 0, vcross(d2, 0, 1), v);
                    sq(dof_at(L, d1, m0) >= 0, vcross(d2, 1, 0), m0);
                    sq(dof_at(L, d1, m0 + unit(d1)) >= 0, vcross(d2, 1, 1), m0);
                    sq(dof_at(L, d2, v) >= 0, vcross(d1, 0, 0), v);
                    sq(dof_at(L, d2, m0) >= 0, vcross(d1, 0, 1), m0);
                    sq(dof_at(L, d2, v + unit(d2)) >= 0, vcross(d1, 1, 0), v);
                    sq(dof_at(L, d2, m0 + unit(d2)) >= 0, vcross(d1, 1, 1), m0);
                    vals[c][i * kLeafSize + off] = den != 0 ? num / den : 0.f;
                }
            }
        });
    }
    int32_t ns = int32_t(L.pool.size() / kLeafSize);
    for (int c = 0; c < 3; ++c)
        for (int32_t i : L.chan_leaves[c]) {
            L.slot[size_t(i) * kVSlots + kVSpai0 + c] = ns++;
            L.pool.insert(L.pool.end(), &vals[c][size_t(i) * kLeafSize], &vals[c][size_t(i + 1) * kLeafSize]);
            L.pool_on.insert(L.pool_on.end(), &L.mask[(size_t(i) * 3 + c) * 8], &L.mask[(size_t(i) * 3 + c) * 8 + 8]);
        }
    L.spai0 = true;
}

void visc_explicit_matrix(const ViscLevelHost& L, std::vector<int64_t>& ptr, std::vector<int32_t>& col,
                          std::vector<float>& val) {
    std::vector<std::vector<std::pair<int32_t, float>>> rows(size_t(L.ndof));
    for (int c = 0; c < 3; ++c) {
        const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        for (int32_t i : L.chan_leaves[c]) {
            const Coord o = L.origin[i];
            for (int off = 0; off < kLeafSize; ++off) {
                if (!L.is_dof(size_t(i), c, off)) continue;
                const Coord v{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                const int32_t me = dof_at(L, c, v);
                auto& row = rows[size_t(me)];
                row.push_back({me, L.coef(kVDiag + c, size_t(i), off)});
                for (int f = 0; f < 6; ++f) {
                    const int a = f / 2;
                    const Coord nb = shifted(v, a, f % 2 == 0 ? 1 : -1);
                    const int32_t nd = dof_at(L, d0, nb);
                    if (nd < 0) continue;
                    row.push_back({nd, f % 2 == 0 ? coef_at(L, vneg(c, a), nb) : L.coef(vneg(c, a), size_t(i), off)});
                }
                const Coord m0 = shifted(v, d0, -1);
                auto add = [&](int ch, const Coord& p, int gid, const Coord& q) {
                    const int32_t nd = dof_at(L, ch, p);
                    if (nd >= 0) row.push_back({nd, coef_at(L, gid, q)});
                };
                add(d1, v, vcross(d2, 0, 0), v);
                add(d1, v + unit(d1), vcross(d2, 0, 1), v);
                add(d1, m0, vcross(d2, 1, 0), m0);
                add(d1, m0 + unit(d1), vcross(d2, 1, 1), m0);
                add(d2, v, vcross(d1, 0, 0), v);
                add(d2, m0, vcross(d1, 0, 1), m0);
                add(d2, v + unit(d2), vcross(d1, 1, 0), v);
                add(d2, m0 + unit(d2), vcross(d1, 1, 1), m0);
            }
        }
    }
    ptr.assign(size_t(L.ndof) + 1, 0);
    col.clear();
    val.clear();
    for (size_t r = 0; r < rows.size(); ++r) {
        auto& row = rows[r];
        std::stable_sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < row.size(); ++k) {
            if (k > 0 && row[k].first == row[k - 1].first) val.back() += row[k].second;
            else {
                col.push_back(row[k].first);
                val.push_back(row[k].second);
            }
        }
        ptr[r + 1] = int64_t(col.size());
    }
}

// ------------------------------------------------------------- rhs
namespace {

// uget(c, v): the liquid velocity of channel c at voxel v as an accessor reads it
template <class UGet>
void build_visc_rhs_impl(const ViscHierarchyHost& h, const UGet& uget, const HostGrid* solid_u, float* out) {
    const ViscGeometry& G = h.geo;
    const ViscLevelHost& L0 = h.levels[0];
    const float eps = 1e-1f;
    LeafSet vls, dls;
    vls.org = G.vol_origin;
    vls.lk.build(vls.org);
    dls.org = G.dof_origin;
    dls.lk.build(dls.org);
    auto fld = [&](const std::vector<float>& v) {
        FieldAcc a;
        a.ls = &vls;
        a.data = v.data();
        a.bg = 0.f;
        return a;
    };
    auto dacc = [&](int c) {
        DofAcc a;
        a.ls = &dls;
        a.data = G.dof_index[c].data();
        a.bg = -1;
        return a;
    };
    auto uval = [&](const HostGrid& g, const Coord& p, int comp) {
        const int32_t i = g.lookup.find(leaf_origin_of(p));
        if (i < 0) return g.background[comp];
        return g.values[(size_t(i) * kLeafSize + off_of(p)) * 3 + comp];
    };
    // first part: u * max(face volume, 0.1) on every DOF
    for (int c = 0; c < 3; ++c) {
        leaf_for(int64_t(G.dof_origin.size()), [&](int64_t b, int64_t e) {
            FieldAcc fa = fld(G.face_vol[c]);
            for (int64_t li = b; li < e; ++li) {
                const Coord o = G.dof_origin[li];
                const int32_t* idx = &G.dof_index[c][size_t(li) * kLeafSize];
                for (int off = 0; off < kLeafSize; ++off) {
                    if (idx[off] < 0) continue;
                    const Coord v{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                    float fv = fa.get(v);
                    if (fv < eps) fv = eps;
                    out[idx[off]] = uget(c, v) * fv;
                }
            }
        });
    }
    (void)L0;
    if (!solid_u) {
        // zero solid velocity: every boundary contribution is +-0, which leaves
        // a nonzero right-hand side unchanged
        return;
    }
    // solid pattern: union of the solid-face masks dilated 3 times (6 neighbours)
    HostGrid uni;
    uni.ncomp = 1;
    uni.origin = G.dof_origin;
    uni.active.assign(G.dof_origin.size() * 8, 0);
    for (int c = 0; c < 3; ++c)
        for (size_t k = 0; k < uni.active.size(); ++k) uni.active[k] |= G.solid_on[c][k];
    const LeafSet pat = dilate(uni, 3, false);
    const float dx = G.dx, factor = G.dt / (G.rho * dx * dx);
    auto solid = [&](int c, const Coord& p) {
        const int32_t i = dls.find(leaf_origin_of(p));
        return i >= 0 && mask_test(&G.solid_on[c][size_t(i) * 8], off_of(p));
    };
    std::vector<std::pair<int64_t, float>> contrib;
    for (int c = 0; c < 3; ++c) {
        const int d0 = c, d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const Coord e0 = unit(d0), e1 = unit(d1), e2 = unit(d2);
        FieldAcc va = fld(G.voxel_vol), ed0 = fld(G.edge_vol[d0]), ed1 = fld(G.edge_vol[d1]), ed2 = fld(G.edge_vol[d2]);
        DofAcc a0 = dacc(d0), a1 = dacc(d1), a2 = dacc(d2);
        GridAcc ma{&G.mu};
        for (const Coord& o : pat.org)
            for (int off = 0; off < kLeafSize; ++off) {
                const Coord v{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
                const int32_t me = a0.get(v);
                float fc[3] = {float(v.x), float(v.y), float(v.z)};
                fc[d0] -= 0.5f;
                auto mus = [&](int a, float da) {
                    float p[3] = {fc[0], fc[1], fc[2]};
                    p[a] += da;
                    return box_sample(ma, p[0], p[1], p[2]);
                };
                if (me != -1) {
                    float t = 0;
                    auto side = [&](const Coord& n, float term) {
                        if (a0.get(n) == -1 && solid(d0, n)) t += term * uval(*solid_u, n, d0);
                    };
                    side(v + e0, 2 * factor * va.get(v) * ma.get(v));
                    {
                        const Coord n = shifted(v, d0, -1);
                        side(n, 2 * factor * va.get(n) * ma.get(n));
                    }
                    side(v + e1, factor * ed2.get(v + e1) * mus(d1, 0.5f));
                    side(shifted(v, d1, -1), factor * ed2.get(v) * mus(d1, -0.5f));
                    side(v + e2, factor * ed1.get(v + e2) * mus(d2, 0.5f));
                    side(shifted(v, d2, -1), factor * ed1.get(v) * mus(d2, -0.5f));
                    if (t != 0) contrib.push_back({me, t});
                }
                const int32_t n1 = a1.get(v), p1 = a1.get(v + e1), n2 = a2.get(v), p2 = a2.get(v + e2);
                const bool sn1 = solid(d1, v), sp1 = solid(d1, v + e1), sn2 = solid(d2, v), sp2 = solid(d2, v + e2);
                auto term = [&](int j, int k) {
                    Coord eo = v;
                    if (j) eo = eo + e1;
                    if (k) eo = eo + e2;
                    float p[3] = {float(v.x), float(v.y), float(v.z)};
                    p[d1] += j ? 0.5f : -0.5f;
                    p[d2] += k ? 0.5f : -0.5f;
                    const float f = (j + k) % 2 ? -factor : factor;
                    return f * ed0.get(eo) * box_sample(ma, p[0], p[1], p[2]);
                };
                if (n1 != -1 || n2 != -1) {
                    const float t = term(0, 0);
                    if (n1 == -1 && sn1) contrib.push_back({n2, -t * uval(*solid_u, v, d1)});
                    if (n2 == -1 && sn2) contrib.push_back({n1, -t * uval(*solid_u, v, d2)});
                }
                if (n1 != -1 || p2 != -1) {
                    const float t = term(0, 1);
                    if (n1 == -1 && sn1) contrib.push_back({p2, -t * uval(*solid_u, v, d1)});
                    if (p2 == -1 && sp2) contrib.push_back({n1, -t * uval(*solid_u, v + e2, d2)});
                }
                if (p1 != -1 || n2 != -1) {
                    const float t = term(1, 0);
                    if (p1 == -1 && sp1) contrib.push_back({n2, -t * uval(*solid_u, v + e1, d1)});
                    if (n2 == -1 && sn2) contrib.push_back({p1, -t * uval(*solid_u, v, d2)});
                }
                if (p1 != -1 || p2 != -1) {
                    const float t = term(1, 1);
                    if (p1 == -1 && sp1) contrib.push_back({p2, -t * uval(*solid_u, v + e1, d1)});
                    if (p2 == -1 && sp2) contrib.push_back({p1, -t * uval(*solid_u, v + e2, d2)});
                }
            }
    }
    std::unordered_map<int64_t, float> acc;
    acc.reserve(contrib.size() * 2 + 1);
    for (const auto& kc : contrib) {
        auto it = acc.find(kc.first);
        if (it != acc.end()) it->second += kc.second;
        else acc.emplace(kc.first, kc.second);
    }
    for (const auto& kv : acc) out[kv.first] = out[kv.first] + kv.second;
}

float grid_value(const HostGrid& g, const Coord& p, int comp) {
    const int32_t i = g.lookup.find(leaf_origin_of(p));
    if (i < 0) return g.background[comp];
    return g.values[(size_t(i) * kLeafSize + size_t(voxel_offset(p.x & 7, p.y & 7, p.z & 7))) * size_t(g.ncomp) +
                    size_t(comp)];
}

}  // namespace

void build_visc_rhs(const ViscHierarchyHost& h, const HostGrid& u, const HostGrid* solid_u, float* out) {
    build_visc_rhs_impl(h, [&](int c, const Coord& v) { return grid_value(u, v, c); }, solid_u, out);
}

void build_visc_rhs(const ViscHierarchyHost& h, const HostGrid* const channels[3], const HostGrid* solid_u,
                    float* out) {
    build_visc_rhs_impl(h, [&](int c, const Coord& v) { return grid_value(*channels[c], v, 0); }, solid_u, out);
}

}  // namespace uaamg
