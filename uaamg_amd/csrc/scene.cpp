#include "scene.h"
#include "fraction_inside.h"

#include <algorithm>
#include <cmath>
#include <unordered_map>

#include "leaf_topology.h"
#include "poisson_host.h"

namespace uaamg {

namespace {
inline uint64_t splitmix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
}  // namespace

float hash_uniform(uint64_t seed, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
    uint64_t h = splitmix(seed);
    h = splitmix(h ^ a);
    h = splitmix(h ^ b);
    h = splitmix(h ^ c);
    h = splitmix(h ^ d);
    return float(h >> 40) * (1.0f / 16777216.0f);
}

float fraction_inside(float phi_left, float phi_right) { return fraction_inside2(phi_left, phi_right); }

float fraction_inside(float phi_bl, float phi_br, float phi_tl, float phi_tr) {
    return fraction_inside4(phi_bl, phi_br, phi_tl, phi_tr);
}

namespace {
struct SdfEval {
    const SceneSpec* s;
    float liquid(float x, float y, float z) const {
        float d = 1e30f;
        if (s->pool > 0.f) d = y - s->pool;
        for (size_t i = 0; i + 3 < s->spheres.size(); i += 4) {
            float dx = x - s->spheres[i], dy = y - s->spheres[i + 1], dz = z - s->spheres[i + 2];
            d = std::min(d, std::sqrt(dx * dx + dy * dy + dz * dz) - s->spheres[i + 3]);
        }
        return d;
    }
    float solid(float x, float y, float z) const {
        float dx = x - s->solid[0], dy = y - s->solid[1], dz = z - s->solid[2];
        return std::sqrt(dx * dx + dy * dy + dz * dz) - s->solid[3];
    }
};
}  // namespace

void build_scene(const SceneSpec& spec, Scene& out) {
    out = Scene();
    out.spec = spec;
    const int n = spec.n;
    const float dx = 1.0f / float(n);
    out.dx = dx;
    const float band = spec.band_voxels * dx;
    out.phi_background = band;
    SdfEval sdf{&spec};
    const int nl = (n + 7) / 8;

    // pass 1: which leaves hold liquid within the band (x-major leaf loop)
    std::vector<uint8_t> keep(size_t(nl) * nl * nl, 0);
    parallel_for(int64_t(nl), [&](int64_t b, int64_t e) {
        for (int64_t lx = b; lx < e; ++lx)
            for (int ly = 0; ly < nl; ++ly)
                for (int lz = 0; lz < nl; ++lz) {
                    bool k = false;
                    for (int i = 0; i < 8 && !k; ++i)
                        for (int j = 0; j < 8 && !k; ++j)
                            for (int l = 0; l < 8 && !k; ++l) {
                                int gx = int(lx) * 8 + i, gy = ly * 8 + j, gz = lz * 8 + l;
                                if (gx >= n || gy >= n || gz >= n) continue;
                                if (sdf.liquid(gx * dx, gy * dx, gz * dx) < band) k = true;
                            }
                    keep[(size_t(lx) * nl + ly) * nl + lz] = k;
                }
    });
    std::vector<Coord> org;
    for (int lx = 0; lx < nl; ++lx)
        for (int ly = 0; ly < nl; ++ly)
            for (int lz = 0; lz < nl; ++lz)
                if (keep[(size_t(lx) * nl + ly) * nl + lz]) org.push_back(Coord{lx * 8, ly * 8, lz * 8});
    std::sort(org.begin(), org.end(), tree_order_less);
    const size_t nleaf = org.size();
    out.origins.resize(nleaf * 3);
    out.active.assign(nleaf * 8, 0);
    out.phi.resize(nleaf * kLeafSize);
    out.weights.resize(nleaf * kLeafSize * 3);
    for (int c = 0; c < 3; ++c) out.vel[c].resize(nleaf * kLeafSize);

    auto in_domain = [n](int x, int y, int z) { return x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n; };
    parallel_for(int64_t(nleaf), [&](int64_t b, int64_t e) {
        for (int64_t li = b; li < e; ++li) {
            const Coord o = org[li];
            out.origins[li * 3 + 0] = o.x;
            out.origins[li * 3 + 1] = o.y;
            out.origins[li * 3 + 2] = o.z;
            for (int off = 0; off < kLeafSize; ++off) {
                const int gx = o.x + (off >> 6), gy = o.y + ((off >> 3) & 7), gz = o.z + (off & 7);
                const size_t idx = size_t(li) * kLeafSize + off;
                const bool inside = in_domain(gx, gy, gz);
                if (inside) {
                    out.active[li * 8 + (off >> 6)] |= uint64_t(1) << (off & 63);
                    out.phi[idx] = sdf.liquid(gx * dx, gy * dx, gz * dx);
                } else {
                    out.phi[idx] = band;
                }
                for (int c = 0; c < 3; ++c) {
                    int px = gx - (c == 0), py = gy - (c == 1), pz = gz - (c == 2);
                    float w = 0.f;
                    if (inside && in_domain(px, py, pz)) {
                        w = 1.f;
                        if (spec.has_solid) {
                            // corners of the face between v-e_c and v, at the voxel's
                            // lower corner plane along c (vertex transform shifted -dx/2)
                            float cx0 = (gx - 0.5f) * dx, cy0 = (gy - 0.5f) * dx, cz0 = (gz - 0.5f) * dx;
                            float s[2][2][2];
                            for (int ii = 0; ii < 2; ++ii)
                                for (int jj = 0; jj < 2; ++jj)
                                    for (int kk = 0; kk < 2; ++kk)
                                        s[ii][jj][kk] = sdf.solid(cx0 + ii * dx, cy0 + jj * dx, cz0 + kk * dx);
                            float frac;
                            if (c == 0) frac = fraction_inside(s[0][0][0], s[0][1][0], s[0][0][1], s[0][1][1]);
                            else if (c == 1) frac = fraction_inside(s[0][0][0], s[0][0][1], s[1][0][0], s[1][0][1]);
                            else frac = fraction_inside(s[0][0][0], s[1][0][0], s[0][1][0], s[1][1][0]);
                            w = std::max(0.f, std::min(1.0f - frac, 1.f));
                        }
                    }
                    out.weights[idx * 3 + c] = w;
                    out.vel[c][idx] = 2.f * hash_uniform(spec.velocity_seed, uint64_t(c) + 1, uint64_t(uint32_t(gx)),
                                                         uint64_t(uint32_t(gy)), uint64_t(uint32_t(gz))) -
                                      1.f;
                }
            }
        }
    });
}

void build_visc_scene(const SceneSpec& spec, float mu, float dt, float rho, ViscScene& out) {
    out = ViscScene();
    build_scene(spec, out.liquid);
    out.mu = mu;
    out.dt = dt;
    out.rho = rho;
    const int n = spec.n;
    const float dx = out.liquid.dx;
    const float band = 3.0f * dx;
    out.solid_background = band;
    SdfEval sdf{&spec};
    auto solid_phi = [&](int ix, int iy, int iz) {
        const int m = std::min(std::min(std::min(ix, n - ix), std::min(iy, n - iy)), std::min(iz, n - iz));
        float d = dx * float(m);
        if (spec.has_solid) d = std::min(d, sdf.solid((ix - 0.5f) * dx, (iy - 0.5f) * dx, (iz - 0.5f) * dx));
        return d;
    };
    // vertex leaves covering [-8, n + 8)
    std::vector<Coord> sorg;
    for (int lx = -8; lx < n + 8; lx += 8)
        for (int ly = -8; ly < n + 8; ly += 8)
            for (int lz = -8; lz < n + 8; lz += 8) sorg.push_back(Coord{lx, ly, lz});
    std::vector<uint8_t> keep(sorg.size(), 0);
    std::vector<float> vals(sorg.size() * kLeafSize);
    std::vector<uint64_t> act(sorg.size() * 8, 0);
    parallel_for(int64_t(sorg.size()), [&](int64_t b, int64_t e) {
        for (int64_t li = b; li < e; ++li) {
            const Coord o = sorg[li];
            for (int off = 0; off < kLeafSize; ++off) {
                const float d = solid_phi(o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7));
                if (std::fabs(d) < band) {
                    keep[li] = 1;
                    act[li * 8 + (off >> 6)] |= uint64_t(1) << (off & 63);
                    vals[li * kLeafSize + off] = d;
                } else {
                    vals[li * kLeafSize + off] = d < 0 ? -band : band;
                }
            }
        }
    });
    for (size_t li = 0; li < sorg.size(); ++li) {
        if (!keep[li]) continue;
        out.solid_origins.insert(out.solid_origins.end(), {sorg[li].x, sorg[li].y, sorg[li].z});
        out.solid_active.insert(out.solid_active.end(), act.begin() + li * 8, act.begin() + li * 8 + 8);
        out.solid_phi.insert(out.solid_phi.end(), vals.begin() + li * kLeafSize, vals.begin() + (li + 1) * kLeafSize);
    }
    // velocity leaves: liquid leaves and their 26 neighbours
    std::vector<Coord> vorg;
    {
        std::unordered_map<uint64_t, int> m;
        const Scene& L = out.liquid;
        for (int64_t i = 0; i < L.nleaf(); ++i)
            for (int ax = -1; ax <= 1; ++ax)
                for (int ay = -1; ay <= 1; ++ay)
                    for (int az = -1; az <= 1; ++az) {
                        const Coord c{L.origins[3 * i] + 8 * ax, L.origins[3 * i + 1] + 8 * ay, L.origins[3 * i + 2] + 8 * az};
                        if (m.emplace(pack_origin(c), 1).second) vorg.push_back(c);
                    }
    }
    std::sort(vorg.begin(), vorg.end(), tree_order_less);
    out.vel_origins.resize(vorg.size() * 3);
    out.vel.resize(vorg.size() * kLeafSize * 3);
    parallel_for(int64_t(vorg.size()), [&](int64_t b, int64_t e) {
        for (int64_t li = b; li < e; ++li) {
            const Coord o = vorg[li];
            out.vel_origins[3 * li] = o.x;
            out.vel_origins[3 * li + 1] = o.y;
            out.vel_origins[3 * li + 2] = o.z;
            for (int off = 0; off < kLeafSize; ++off) {
                const int gx = o.x + (off >> 6), gy = o.y + ((off >> 3) & 7), gz = o.z + (off & 7);
                for (int c = 0; c < 3; ++c)
                    out.vel[(size_t(li) * kLeafSize + off) * 3 + c] =
                        2.f * hash_uniform(spec.velocity_seed, uint64_t(c) + 11, uint64_t(uint32_t(gx)),
                                           uint64_t(uint32_t(gy)), uint64_t(uint32_t(gz))) -
                        1.f;
            }
        }
    });
}

bool scene_preset(int config, int32_t n, uint64_t seed, SceneSpec& s) {
    s = SceneSpec();
    s.velocity_seed = seed;
    switch (config) {
        case 0:
            s.n = n > 0 ? n : 128;
            s.pool = 0.4f;
            s.spheres = {0.5f, 0.7f, 0.5f, 0.15f};
            s.has_solid = false;
            return true;
        case 1: {
            s.n = n > 0 ? n : 256;
            s.pool = 0.3f;
            s.has_solid = true;
            s.solid[0] = s.solid[1] = s.solid[2] = 0.5f;
            s.solid[3] = 0.2f;
            for (int i = 0; i < 8; ++i) {
                for (int c = 0; c < 3; ++c) s.spheres.push_back(hash_uniform(seed, 101, uint64_t(i), uint64_t(c), 0));
                s.spheres.push_back(0.05f + 0.15f * hash_uniform(seed, 102, uint64_t(i), 0, 0));
            }
            return true;
        }
        case 3: {
            float u = hash_uniform(seed, 301, 0, 0, 0);
            int edge = int(std::lround(32.0 * std::exp(double(u) * std::log(192.0 / 32.0))));
            s.n = n > 0 ? n : std::max(32, std::min(192, edge));
            s.pool = 0.1f + 0.4f * hash_uniform(seed, 302, 0, 0, 0);
            int ns = 1 + std::min(5, int(6.f * hash_uniform(seed, 303, 0, 0, 0)));
            for (int i = 0; i < ns; ++i) {
                for (int c = 0; c < 3; ++c) s.spheres.push_back(hash_uniform(seed, 304, uint64_t(i), uint64_t(c), 0));
                s.spheres.push_back(0.05f + 0.15f * hash_uniform(seed, 305, uint64_t(i), 0, 0));
            }
            s.has_solid = hash_uniform(seed, 306, 0, 0, 0) < 0.5f;
            for (int c = 0; c < 3; ++c) s.solid[c] = 0.3f + 0.4f * hash_uniform(seed, 307, uint64_t(c), 0, 0);
            s.solid[3] = 0.1f + 0.1f * hash_uniform(seed, 308, 0, 0, 0);
            return true;
        }
        default: return false;
    }
}

}  // namespace uaamg
