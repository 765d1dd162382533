// ORACLE - TEST INFRASTRUCTURE ONLY.
//
// OpenVDB-style sparse grids for the CPU restatements: 8^3 leaves keyed by
// origin in Tree4<5,4,3> traversal order, read through accessor-style get()
// calls (stored value inside a leaf, active or not; background outside).
#pragma once

#include <cstdint>
#include <cstring>
#include <map>

namespace orc {

struct Coord {
    int32_t x, y, z;
    int32_t& operator[](int c) { return c == 0 ? x : (c == 1 ? y : z); }
    int32_t operator[](int c) const { return c == 0 ? x : (c == 1 ? y : z); }
    bool operator==(const Coord& o) const { return x == o.x && y == o.y && z == o.z; }
};

inline Coord leaf_of(Coord c) { return {c.x & ~7, c.y & ~7, c.z & ~7}; }
inline int offset_of(Coord c) { return ((c.x & 7) << 6) | ((c.y & 7) << 3) | (c.z & 7); }
inline Coord local_of(int off) { return {off >> 6, (off >> 3) & 7, off & 7}; }

// OpenVDB Tree4<5,4,3>: root table sorted by Coord (lexicographic, signed) of
// the 4096-block origin, then upper-node child index (x-major, 32^3 of 128),
// then lower-node child index (x-major, 16^3 of 8).
struct TreeLess {
    bool operator()(const Coord& a, const Coord& b) const {
        const int32_t ra[3] = {a.x & ~4095, a.y & ~4095, a.z & ~4095};
        const int32_t rb[3] = {b.x & ~4095, b.y & ~4095, b.z & ~4095};
        for (int i = 0; i < 3; ++i)
            if (ra[i] != rb[i]) return ra[i] < rb[i];
        auto upper = [](const Coord& c) {
            return (((c.x & 4095) >> 7) << 10) | (((c.y & 4095) >> 7) << 5) | ((c.z & 4095) >> 7);
        };
        auto lower = [](const Coord& c) { return (((c.x & 127) >> 3) << 8) | (((c.y & 127) >> 3) << 4) | ((c.z & 127) >> 3); };
        if (upper(a) != upper(b)) return upper(a) < upper(b);
        return lower(a) < lower(b);
    }
};

template <class T>
struct Leaf {
    uint64_t on[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    T v[512];
    bool is_on(int off) const { return (on[off >> 6] >> (off & 63)) & 1u; }
    void set_on(int off, bool b) {
        if (b) on[off >> 6] |= uint64_t(1) << (off & 63);
        else on[off >> 6] &= ~(uint64_t(1) << (off & 63));
    }
    int count() const {
        int n = 0;
        for (int w = 0; w < 8; ++w) n += __builtin_popcountll(on[w]);
        return n;
    }
};

template <class T>
struct Grid {
    std::map<Coord, Leaf<T>, TreeLess> leaves;
    T background{};
    const Leaf<T>* probe(Coord origin) const {
        auto it = leaves.find(origin);
        return it == leaves.end() ? nullptr : &it->second;
    }
    Leaf<T>* probe(Coord origin) {
        auto it = leaves.find(origin);
        return it == leaves.end() ? nullptr : &it->second;
    }
    T get(Coord c) const {
        const Leaf<T>* l = probe(leaf_of(c));
        return l ? l->v[offset_of(c)] : background;
    }
    bool is_on(Coord c) const {
        const Leaf<T>* l = probe(leaf_of(c));
        return l && l->is_on(offset_of(c));
    }
    // TopologyCopy: same leaves and active masks, every value = fill
    template <class U>
    void topology_copy(const Grid<U>& src, T fill) {
        leaves.clear();
        background = fill;
        for (const auto& kv : src.leaves) {
            Leaf<T>& l = leaves[kv.first];
            std::memcpy(l.on, kv.second.on, sizeof(l.on));
            for (int i = 0; i < 512; ++i) l.v[i] = fill;
        }
    }
};

struct Vec3 {
    float v[3];
};

using FGrid = Grid<float>;
using IGrid = Grid<int32_t>;
using VGrid = Grid<Vec3>;

}  // namespace orc
