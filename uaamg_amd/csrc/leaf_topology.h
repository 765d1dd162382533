// Sparse 8^3 leaf-brick topology on the host.
//
// A grid is a set of 8x8x8 leaves, each with an origin that is a multiple of 8,
// a 512-bit active mask and 512 values per component. Inside a leaf the voxel
// offset is 64*x + 8*y + z (z fastest), which is the layout of OpenVDB's
// LeafNode<T,3> buffer (simd_vdb_poisson.cpp:750-751 relies on it).
//
// Leaves are kept in the traversal order of OpenVDB's Tree4<T,5,4,3>: root
// children (4096^3 blocks) sorted lexicographically by their signed origin,
// then the 32^3 child index of the upper internal node (x-major), then the
// 16^3 child index of the lower internal node (x-major). LeafManager visits
// leaves in this order, and the DOF numbering of setDofIndex
// (simd_vdb_poisson.cpp:637-671) is a prefix scan in this order.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <utility>
#include <vector>

namespace uaamg {

constexpr int kLeafDim = 8;
constexpr int kLeafSize = 512;

struct Coord {
    int32_t x, y, z;
    Coord operator+(const Coord& o) const { return {x + o.x, y + o.y, z + o.z}; }
    bool operator==(const Coord& o) const { return x == o.x && y == o.y && z == o.z; }
    int32_t operator[](int c) const { return c == 0 ? x : (c == 1 ? y : z); }
    int32_t& operator[](int c) { return c == 0 ? x : (c == 1 ? y : z); }
};

// floor division by 8 / origin of the leaf containing a voxel
inline int32_t leaf_floor(int32_t v) { return v & ~int32_t(7); }
inline Coord leaf_origin_of(const Coord& v) { return {leaf_floor(v.x), leaf_floor(v.y), leaf_floor(v.z)}; }
inline int voxel_offset(int lx, int ly, int lz) { return (lx << 6) | (ly << 3) | lz; }

// Sort key reproducing Tree4<T,5,4,3> leaf traversal order.
struct TreeOrderKey {
    int32_t rx, ry, rz;   // root child origin (multiple of 4096, signed)
    int32_t upper;        // child index inside the 32^3 internal node
    int32_t lower;        // child index inside the 16^3 internal node
    bool operator<(const TreeOrderKey& o) const {
        if (rx != o.rx) return rx < o.rx;
        if (ry != o.ry) return ry < o.ry;
        if (rz != o.rz) return rz < o.rz;
        if (upper != o.upper) return upper < o.upper;
        return lower < o.lower;
    }
};

inline TreeOrderKey tree_order_key(const Coord& origin) {
    TreeOrderKey k;
    k.rx = origin.x & ~int32_t(4095);
    k.ry = origin.y & ~int32_t(4095);
    k.rz = origin.z & ~int32_t(4095);
    k.upper = (((origin.x & 4095) >> 7) << 10) | (((origin.y & 4095) >> 7) << 5) | ((origin.z & 4095) >> 7);
    k.lower = (((origin.x & 127) >> 3) << 8) | (((origin.y & 127) >> 3) << 4) | ((origin.z & 127) >> 3);
    return k;
}

inline bool tree_order_less(const Coord& a, const Coord& b) { return tree_order_key(a) < tree_order_key(b); }

inline uint64_t pack_origin(const Coord& o) {
    // origins are multiples of 8; 21 bits per axis after the shift covers +-2^23 voxels
    const uint64_t m = (uint64_t(1) << 21) - 1;
    return ((uint64_t(uint32_t(o.x >> 3)) & m) << 42) | ((uint64_t(uint32_t(o.y >> 3)) & m) << 21) |
           (uint64_t(uint32_t(o.z >> 3)) & m);
}

// Allocator whose resize() leaves new elements uninitialised (value-less
// construction), for large bricks that are written in full right afterwards:
// it skips a single-threaded zero fill of every page.
template <class T>
struct DefaultInitAllocator : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = DefaultInitAllocator<U>;
    };
    DefaultInitAllocator() = default;
    template <class U>
    DefaultInitAllocator(const DefaultInitAllocator<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... Args>
    void construct(U* p, Args&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
};
using BrickVec = std::vector<float, DefaultInitAllocator<float>>;

// Origin -> position lookup for a leaf list: an open-addressing table
// (linear probing, multiplicative hash of the packed origin, load <= 1/2).
// The first of duplicate origins wins.
class LeafLookup {
public:
    void build(const std::vector<Coord>& origins) {
        int bits = 4;
        while ((size_t(1) << bits) < origins.size() * 2) ++bits;
        shift_ = 64 - bits;
        mask_ = (size_t(1) << bits) - 1;
        keys_.assign(mask_ + 1, kEmpty);
        vals_.assign(mask_ + 1, -1);
        for (size_t i = 0; i < origins.size(); ++i) {
            const uint64_t k = pack_origin(origins[i]);
            size_t h = slot(k);
            while (keys_[h] != kEmpty && keys_[h] != k) h = (h + 1) & mask_;
            if (keys_[h] == kEmpty) {
                keys_[h] = k;
                vals_[h] = int32_t(i);
            }
        }
    }
    int32_t find(const Coord& origin) const {
        if (keys_.empty()) return -1;
        const uint64_t k = pack_origin(origin);
        for (size_t h = slot(k);; h = (h + 1) & mask_) {
            const uint64_t q = keys_[h];
            if (q == k) return vals_[h];
            if (q == kEmpty) return -1;
        }
    }

private:
    // packed origins use 63 bits, so the all-ones word marks an empty slot
    static constexpr uint64_t kEmpty = ~uint64_t(0);
    size_t slot(uint64_t k) const { return size_t((k * 0x9E3779B97F4A7C15ull) >> shift_); }
    std::vector<uint64_t> keys_;
    std::vector<int32_t> vals_;
    size_t mask_ = 0;
    int shift_ = 60;
};

// Packed origin -> int32 map with insertion (open addressing, linear probing,
// grows at load 1/2); emplace keeps the first value of a key.
class OriginMap {
public:
    // (value, inserted)
    std::pair<int32_t, bool> emplace(uint64_t k, int32_t v) {
        if ((count_ + 1) * 2 > keys_.size()) grow();
        size_t h = slot(k);
        while (keys_[h] != kEmpty) {
            if (keys_[h] == k) {
                if (vals_[h] >= 0) return {vals_[h], false};
                vals_[h] = v;   // a removed key comes back
                return {v, true};
            }
            h = (h + 1) & mask_;
        }
        keys_[h] = k;
        vals_[h] = v;
        ++count_;
        return {v, true};
    }
    int32_t find(uint64_t k) const {
        if (keys_.empty()) return -1;
        for (size_t h = slot(k);; h = (h + 1) & mask_) {
            const uint64_t q = keys_[h];
            if (q == k) return vals_[h];
            if (q == kEmpty) return -1;
        }
    }
    // values are >= 0; a removed key keeps its slot (so probes pass it) with value -1
    void remove(uint64_t k) {
        if (keys_.empty()) return;
        for (size_t h = slot(k);; h = (h + 1) & mask_) {
            if (keys_[h] == k) {
                vals_[h] = -1;
                return;
            }
            if (keys_[h] == kEmpty) return;
        }
    }
    // f(key, value) for every present key, in slot order
    template <class F>
    void for_each(F&& f) const {
        for (size_t h = 0; h < keys_.size(); ++h)
            if (keys_[h] != kEmpty && vals_[h] >= 0) f(keys_[h], vals_[h]);
    }

private:
    static constexpr uint64_t kEmpty = ~uint64_t(0);
    size_t slot(uint64_t k) const { return size_t((k * 0x9E3779B97F4A7C15ull) >> shift_); }
    void grow() {
        std::vector<uint64_t> ok;
        std::vector<int32_t> ov;
        ok.swap(keys_);
        ov.swap(vals_);
        const int bits = ok.empty() ? 4 : 64 - shift_ + 1;
        shift_ = 64 - bits;
        mask_ = (size_t(1) << bits) - 1;
        keys_.assign(mask_ + 1, kEmpty);
        vals_.assign(mask_ + 1, -1);
        for (size_t i = 0; i < ok.size(); ++i) {
            if (ok[i] == kEmpty) continue;
            size_t h = slot(ok[i]);
            while (keys_[h] != kEmpty) h = (h + 1) & mask_;
            keys_[h] = ok[i];
            vals_[h] = ov[i];
        }
    }
    std::vector<uint64_t> keys_;
    std::vector<int32_t> vals_;
    size_t count_ = 0, mask_ = 0;
    int shift_ = 60;
};

inline bool mask_test(const uint64_t* mask8, int offset) { return (mask8[offset >> 6] >> (offset & 63)) & 1u; }
inline void mask_set(uint64_t* mask8, int offset) { mask8[offset >> 6] |= uint64_t(1) << (offset & 63); }
inline int mask_count(const uint64_t* mask8) {
    int n = 0;
    for (int w = 0; w < 8; ++w) n += __builtin_popcountll(mask8[w]);
    return n;
}

// A host copy of one flattened sparse grid (FloatGrid or Vec3fGrid), leaves in
// tree order. value() follows ValueAccessor::getValue: the stored value when
// the leaf exists (active or not), otherwise the background.
struct HostGrid {
    int ncomp = 1;
    float background[3] = {0.f, 0.f, 0.f};
    std::vector<Coord> origin;
    std::vector<uint64_t> active;   // [nleaf][8]
    // [nleaf][512][ncomp] (Vec3f leaves are AoS): the caller's array read in
    // place when make_host_grid may borrow it and its leaves are already in
    // tree order, otherwise the grid's own copy in `owned`
    const float* values = nullptr;
    std::unique_ptr<float[]> owned;
    LeafLookup lookup;

    HostGrid() = default;
    HostGrid(HostGrid&&) = default;   // a moved vector keeps its buffer, so values stays valid
    HostGrid& operator=(HostGrid&&) = default;
    // a copy always owns its values (it may outlive a borrowed array)
    HostGrid(const HostGrid& o) { *this = o; }
    HostGrid& operator=(const HostGrid& o) {
        if (this == &o) return *this;
        ncomp = o.ncomp;
        for (int c = 0; c < 3; ++c) background[c] = o.background[c];
        origin = o.origin;
        active = o.active;
        lookup = o.lookup;
        const size_t n = o.values ? o.origin.size() * kLeafSize * size_t(o.ncomp) : 0;
        owned.reset(n ? new float[n] : nullptr);
        if (n) std::copy(o.values, o.values + n, owned.get());
        values = owned.get();
        return *this;
    }

    size_t nleaf() const { return origin.size(); }
    const float* leaf_values(size_t i) const { return values + i * kLeafSize * ncomp; }

    float value(const Coord& v, int comp = 0) const {
        int32_t li = lookup.find(leaf_origin_of(v));
        if (li < 0) return background[comp];
        int off = voxel_offset(v.x & 7, v.y & 7, v.z & 7);
        return values[(size_t(li) * kLeafSize + off) * ncomp + comp];
    }
};

// Build a tree-ordered HostGrid from caller arrays in any leaf order. With
// borrow set and the leaves already in tree order the values are read in
// place (the caller keeps them alive while the grid is used); otherwise they
// are copied. Returns false on malformed input (origin not a multiple of 8,
// duplicates).
bool make_host_grid(int64_t nleaf, const int32_t* origins, const uint64_t* active, const float* values,
                    int ncomp, const float* background, HostGrid& out, bool borrow = false);

}  // namespace uaamg
