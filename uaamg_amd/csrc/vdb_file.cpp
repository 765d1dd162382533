// OpenVDB file reader (see vdb_file.h).
#include "vdb_file.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <tuple>

namespace uaamg {

// ------------------------------------------------------------------ LZ4 block
// LZ4 block format: sequences of [token][literal length ext][literals]
// [offset LE16][match length ext]; token high nibble = literal count, low
// nibble = match length - 4, a nibble of 15 continues in bytes of 255; the
// last sequence carries literals only.
bool lz4_block_decompress(const uint8_t* src, size_t srclen, uint8_t* dst, size_t dstlen, size_t* written) {
    size_t ip = 0, op = 0;
    while (ip < srclen) {
        const uint8_t token = src[ip++];
        size_t lit = token >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= srclen) return false;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit > srclen - ip || lit > dstlen - op) return false;
        std::memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip == srclen) break;   // the last sequence: literals only
        if (srclen - ip < 2) return false;
        const size_t offset = size_t(src[ip]) | (size_t(src[ip + 1]) << 8);
        ip += 2;
        if (offset == 0 || offset > op) return false;
        size_t ml = token & 15u;
        if (ml == 15) {
            uint8_t b;
            do {
                if (ip >= srclen) return false;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        ml += 4;
        if (ml > dstlen - op) return false;
        // the match may overlap its own output: copy byte by byte
        const uint8_t* m = dst + op - offset;
        for (size_t k = 0; k < ml; ++k) dst[op + k] = m[k];
        op += ml;
    }
    if (written) *written = op;
    return true;
}

// --------------------------------------------------------------------- Blosc
// Blosc 1 container: a 16-byte header (format version, codec version, flags,
// type size, uncompressed bytes, block size, total compressed bytes). Flags:
// bit 0 byte shuffle, bit 1 stored uncompressed, bit 2 bit shuffle, bit 4 no
// split, bits 5-7 the codec (0 BloscLZ, 1 LZ4, 2 Snappy, 3 zlib, 4 Zstd).
// Compressed data: the start offset of every block, then per block its splits
// (the block cut into typesize streams when splitting applies: no no-split
// flag, type size <= 16, at least 128 elements, not the short last block),
// each an int32 compressed size and the codec stream (a size equal to the
// split's own means it is stored as is). Shuffle is undone per block.
namespace {
inline uint32_t rd32(const uint8_t* p) { return uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24; }
}  // namespace

bool blosc_decompress(const uint8_t* src, size_t srclen, uint8_t* dst, size_t dstlen, std::string& err) {
    if (srclen < 16) {
        err = "blosc: short header";
        return false;
    }
    const uint8_t version = src[0], flags = src[2], typesize = src[3];
    const uint32_t nbytes = rd32(src + 4), blocksize = rd32(src + 8), cbytes = rd32(src + 12);
    if (version == 0 || version > 3) {
        err = "blosc: unsupported format version " + std::to_string(version);
        return false;
    }
    if (nbytes != dstlen) {
        err = "blosc: uncompressed size " + std::to_string(nbytes) + " where " + std::to_string(dstlen) + " expected";
        return false;
    }
    if (cbytes > srclen) {
        err = "blosc: truncated buffer";
        return false;
    }
    if (nbytes == 0) return true;
    if (flags & 0x2) {   // stored uncompressed
        if (size_t(cbytes) < 16 + size_t(nbytes)) {
            err = "blosc: truncated uncompressed buffer";
            return false;
        }
        std::memcpy(dst, src + 16, nbytes);
        return true;
    }
    if (flags & 0x4) {
        err = "blosc: bit shuffle is not supported";
        return false;
    }
    const int codec = flags >> 5;
    if (codec != 1 && codec != 3) {
        err = "blosc: codec " + std::to_string(codec) + " is not supported (LZ4 and zlib are)";
        return false;
    }
    if (blocksize == 0 || typesize == 0) {
        err = "blosc: zero block or type size";
        return false;
    }
    const bool shuffle = (flags & 0x1) && typesize > 1;
    const bool dont_split = flags & 0x10;
    const size_t nblocks = (size_t(nbytes) + blocksize - 1) / blocksize;
    const size_t leftover = nbytes % blocksize;
    if (16 + 4 * nblocks > cbytes) {
        err = "blosc: truncated block table";
        return false;
    }
    std::vector<uint8_t> tmp(blocksize);
    for (size_t j = 0; j < nblocks; ++j) {
        const bool last_short = (j + 1 == nblocks) && leftover;
        const size_t bsize = last_short ? leftover : blocksize;
        const bool split = !dont_split && typesize <= 16 && bsize / typesize >= 128 && bsize % typesize == 0 && !last_short;
        const size_t nsplits = split ? typesize : 1;
        const size_t neblock = bsize / nsplits;
        size_t pos = rd32(src + 16 + 4 * j);
        uint8_t* out = shuffle ? tmp.data() : dst + j * size_t(blocksize);
        for (size_t k = 0; k < nsplits; ++k) {
            if (pos + 4 > cbytes) {
                err = "blosc: truncated split header";
                return false;
            }
            const size_t csize = rd32(src + pos);
            pos += 4;
            if (csize > cbytes - pos) {
                err = "blosc: truncated split";
                return false;
            }
            uint8_t* o = out + k * neblock;
            if (csize == neblock) {
                std::memcpy(o, src + pos, neblock);
            } else if (codec == 1) {
                size_t got = 0;
                if (!lz4_block_decompress(src + pos, csize, o, neblock, &got) || got != neblock) {
                    err = "blosc: malformed LZ4 stream";
                    return false;
                }
            } else {
                uLongf got = uLongf(neblock);
                if (uncompress(o, &got, src + pos, uLong(csize)) != Z_OK || got != neblock) {
                    err = "blosc: malformed zlib stream";
                    return false;
                }
            }
            pos += csize;
        }
        if (shuffle) {
            uint8_t* d = dst + j * size_t(blocksize);
            const size_t nelem = bsize / typesize;
            for (size_t b = 0; b < typesize; ++b)
                for (size_t i = 0; i < nelem; ++i) d[i * typesize + b] = tmp[b * nelem + i];
            for (size_t i = nelem * typesize; i < bsize; ++i) d[i] = tmp[i];
        }
    }
    return true;
}

// -------------------------------------------------------------- file reading
namespace {

constexpr uint32_t kCompressZip = 0x1, kCompressActiveMask = 0x2, kCompressBlosc = 0x4;
// per-node metadata of the compressed value arrays
enum : int8_t {
    kNoMaskOrInactiveVals = 0,
    kNoMaskAndMinusBg = 1,
    kNoMaskAndOneInactiveVal = 2,
    kMaskAndNoInactiveVals = 3,
    kMaskAndOneInactiveVal = 4,
    kMaskAndTwoInactiveVals = 5,
    kNoMaskAndAllVals = 6,
};

struct Cursor {
    const uint8_t* base;
    size_t size, pos = 0;
    std::string* err;
    bool fail(const std::string& m) {
        if (err->empty()) *err = m + " at byte " + std::to_string(pos);
        return false;
    }
    bool need(size_t n) { return n <= size - std::min(pos, size) ? true : fail("unexpected end of file"); }
    bool bytes(void* out, size_t n) {
        if (!need(n)) return false;
        if (n) std::memcpy(out, base + pos, n);
        pos += n;
        return true;
    }
    template <class T>
    bool get(T& v) { return bytes(&v, sizeof(T)); }
    bool str(std::string& s) {
        uint32_t n;
        if (!get(n) || !need(n)) return false;
        s.assign(reinterpret_cast<const char*>(base + pos), n);
        pos += n;
        return true;
    }
    bool skip(size_t n) {
        if (!need(n)) return false;
        pos += n;
        return true;
    }
};

// IEEE half to float
float half_to_float(uint16_t h) {
    const uint32_t s = uint32_t(h >> 15) << 31;
    uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) bits = s;
        else {   // subnormal: normalise
            int k = -1;
            do {
                ++k;
                m <<= 1;
            } while (!(m & 0x400));
            bits = s | (uint32_t(127 - 15 - k) << 23) | ((m & 0x3ff) << 13);
        }
    } else if (e == 31) {
        bits = s | 0x7f800000u | (m << 13);
    } else {
        bits = s | ((e + 127 - 15) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

struct Mask {
    std::vector<uint64_t> w;
    explicit Mask(size_t nbits = 0) : w((nbits + 63) / 64, 0) {}
    bool on(size_t i) const { return (w[i >> 6] >> (i & 63)) & 1u; }
    size_t count() const {
        size_t c = 0;
        for (uint64_t x : w) c += size_t(__builtin_popcountll(x));
        return c;
    }
    bool load(Cursor& c) { return c.bytes(w.data(), w.size() * 8); }
};

class GridReader {
public:
    GridReader(Cursor& c, VdbGrid& g, uint32_t compression, uint32_t version)
        : c_(c), g_(g), comp_(compression), version_(version) {}

    // the root's topology, then (at block_pos) the leaf buffers
    bool read(size_t block_pos) {
        int32_t buffer_count;
        if (!c_.get(buffer_count)) return false;
        if (buffer_count != 1) return c_.fail("unsupported buffer count " + std::to_string(buffer_count));
        if (!value(bg_)) return false;
        for (int k = 0; k < 3; ++k) g_.background[k] = k < g_.ncomp ? bg_[k] : 0.f;
        uint32_t ntiles, nchildren;
        if (!c_.get(ntiles) || !c_.get(nchildren)) return false;
        for (uint32_t t = 0; t < ntiles; ++t) {
            int32_t o[3];
            float v[3];
            uint8_t act;
            if (!c_.bytes(o, 12) || !value(v) || !c_.get(act)) return false;
            add_tile(o, 4096, v, act != 0);
        }
        for (uint32_t n = 0; n < nchildren; ++n) {
            int32_t o[3];
            if (!c_.bytes(o, 12)) return false;
            if (!internal_topology(o, 5)) return false;
        }
        if (block_pos) c_.pos = block_pos;
        // buffers in the same order as the leaves
        const size_t n = leaf_masks_.size();
        g_.active.resize(n * 8);
        g_.values.assign(n * 512 * size_t(g_.ncomp), 0.f);
        for (size_t i = 0; i < n; ++i) {
            Mask m(512);
            if (!m.load(c_)) return false;
            if (m.w != leaf_masks_[i].w) return c_.fail("leaf buffer mask differs from its topology");
            std::copy(m.w.begin(), m.w.end(), g_.active.begin() + long(i * 8));
            if (!compressed_values(&g_.values[i * 512 * size_t(g_.ncomp)], 512, m)) return false;
        }
        return true;
    }

private:
    // one value of the grid's type at full precision
    bool value(float v[3]) {
        for (int k = 0; k < g_.ncomp; ++k)
            if (!c_.get(v[k])) return false;
        return true;
    }

    void add_tile(const int32_t o[3], int32_t dim, const float v[3], bool act) {
        bool is_bg = !act;
        for (int k = 0; k < g_.ncomp; ++k) is_bg = is_bg && v[k] == bg_[k];
        if (is_bg) return;
        VdbTile t{};
        std::copy(o, o + 3, t.origin);
        t.dim = dim;
        for (int k = 0; k < 3; ++k) t.value[k] = k < g_.ncomp ? v[k] : 0.f;
        t.active = act;
        g_.tiles.push_back(t);
    }

    // InternalNode<child, log2dim>::readTopology: child mask, value mask, the
    // compressed tile values (all entries), then the children in index order
    bool internal_topology(const int32_t origin[3], int log2dim) {
        const size_t nvalues = size_t(1) << (3 * log2dim);
        const int child_log2 = log2dim == 5 ? 7 : 3;   // total log2 edge of a child (internal 4: 128, leaf: 8)
        Mask child(nvalues), vmask(nvalues);
        if (!child.load(c_) || !vmask.load(c_)) return false;
        std::vector<float> vals(nvalues * size_t(g_.ncomp));
        if (!compressed_values(vals.data(), nvalues, vmask)) return false;
        const size_t dim = size_t(1) << log2dim;
        auto child_origin = [&](size_t i, int32_t o[3]) {
            o[0] = origin[0] + int32_t(((i >> (2 * log2dim)) & (dim - 1)) << child_log2);
            o[1] = origin[1] + int32_t(((i >> log2dim) & (dim - 1)) << child_log2);
            o[2] = origin[2] + int32_t((i & (dim - 1)) << child_log2);
        };
        for (size_t i = 0; i < nvalues; ++i) {
            if (child.on(i)) continue;
            int32_t o[3];
            child_origin(i, o);
            add_tile(o, int32_t(1) << child_log2, &vals[i * size_t(g_.ncomp)], vmask.on(i));
        }
        for (size_t i = 0; i < nvalues; ++i) {
            if (!child.on(i)) continue;
            int32_t o[3];
            child_origin(i, o);
            if (log2dim == 5) {
                if (!internal_topology(o, 4)) return false;
            } else {
                Mask m(512);
                if (!m.load(c_)) return false;
                g_.origins.insert(g_.origins.end(), o, o + 3);
                leaf_masks_.push_back(std::move(m));
            }
        }
        return true;
    }

    // readCompressedValues: metadata byte, inactive values, selection mask,
    // then the (active, when mask-compressed) values through the codec
    bool compressed_values(float* dest, size_t count, const Mask& vmask) {
        const int nc = g_.ncomp;
        int8_t meta = kNoMaskAndAllVals;
        if (!c_.get(meta)) return false;
        if (meta < 0 || meta > kNoMaskAndAllVals) return c_.fail("bad value-compression metadata");
        float in0[3], in1[3];
        for (int k = 0; k < 3; ++k) {
            in1[k] = bg_[k];
            in0[k] = meta == kNoMaskOrInactiveVals ? bg_[k] : -bg_[k];
        }
        if (meta == kNoMaskAndOneInactiveVal || meta == kMaskAndOneInactiveVal || meta == kMaskAndTwoInactiveVals) {
            if (!value(in0)) return false;
            if (meta == kMaskAndTwoInactiveVals && !value(in1)) return false;
        }
        Mask sel(count);
        if (meta == kMaskAndNoInactiveVals || meta == kMaskAndOneInactiveVal || meta == kMaskAndTwoInactiveVals)
            if (!sel.load(c_)) return false;
        size_t nread = count;
        const bool packed = (comp_ & kCompressActiveMask) && meta != kNoMaskAndAllVals;
        if (packed) nread = vmask.count();
        std::vector<float> tmp;
        float* buf = dest;
        if (nread != count) {
            tmp.resize(nread * size_t(nc));
            buf = tmp.data();
        }
        if (!read_data(buf, nread * size_t(nc))) return false;
        if (nread != count) {
            size_t t = 0;
            for (size_t i = 0; i < count; ++i) {
                const float* src = vmask.on(i) ? &tmp[(t++) * size_t(nc)] : (sel.on(i) ? in1 : in0);
                for (int k = 0; k < nc; ++k) dest[i * size_t(nc) + size_t(k)] = src[k];
            }
        }
        return true;
    }

    // n floats (or halves widened to floats) through the grid's codec
    bool read_data(float* out, size_t n) {
        const size_t esz = g_.half_float ? 2 : 4;
        const size_t nbytes = n * esz;
        std::vector<uint8_t> raw;
        uint8_t* dst = g_.half_float ? nullptr : reinterpret_cast<uint8_t*>(out);
        if (g_.half_float) {
            raw.resize(nbytes);
            dst = raw.data();
        }
        if (comp_ & (kCompressBlosc | kCompressZip)) {
            int64_t csize;
            if (!c_.get(csize)) return false;
            if (csize <= 0) {   // stored uncompressed
                if (size_t(-csize) != nbytes) return c_.fail("uncompressed block size mismatch");
                if (!c_.bytes(dst, nbytes)) return false;
            } else {
                if (!c_.need(size_t(csize))) return false;
                const uint8_t* src = c_.base + c_.pos;
                if (comp_ & kCompressBlosc) {
                    std::string e;
                    if (!blosc_decompress(src, size_t(csize), dst, nbytes, e)) return c_.fail(e);
                } else {
                    uLongf got = uLongf(nbytes);
                    if (uncompress(dst, &got, src, uLong(csize)) != Z_OK || got != nbytes)
                        return c_.fail("malformed zlib block");
                }
                c_.pos += size_t(csize);
            }
        } else if (!c_.bytes(dst, nbytes)) {
            return false;
        }
        if (g_.half_float)
            for (size_t i = 0; i < n; ++i) {
                uint16_t h;
                std::memcpy(&h, &raw[i * 2], 2);
                out[i] = half_to_float(h);
            }
        return true;
    }

    Cursor& c_;
    VdbGrid& g_;
    uint32_t comp_, version_;
    float bg_[3] = {0, 0, 0};
    std::vector<Mask> leaf_masks_;
};

bool read_meta(Cursor& c, VdbGrid* g) {
    uint32_t n;
    if (!c.get(n)) return false;
    for (uint32_t i = 0; i < n; ++i) {
        std::string name, type;
        uint32_t size;
        if (!c.str(name) || !c.str(type) || !c.get(size) || !c.need(size)) return false;
        const uint8_t* v = c.base + c.pos;
        if (g) {
            if (name == "class" && type == "string") g->grid_class.assign(reinterpret_cast<const char*>(v), size);
            if (name == "file_voxel_count" && type == "int64" && size == 8) std::memcpy(&g->meta_voxel_count, v, 8);
            if (name == "file_bbox_min" && type == "vec3i" && size == 12) {
                std::memcpy(g->meta_bbox, v, 12);
                g->has_meta_bbox = true;
            }
            if (name == "file_bbox_max" && type == "vec3i" && size == 12) std::memcpy(g->meta_bbox + 3, v, 12);
        }
        c.pos += size;
    }
    return true;
}

// Transform::read: the map type name, then the map's members
bool read_transform(Cursor& c, VdbGrid& g) {
    if (!c.str(g.map_type)) return false;
    auto vec3 = [&](double v[3]) { return c.bytes(v, 24); };
    double t[3] = {0, 0, 0}, scale[3], vs[3], tmp[3];
    const std::string& m = g.map_type;
    if (m == "UniformScaleMap" || m == "ScaleMap") {
        if (!vec3(scale) || !vec3(vs) || !vec3(tmp) || !vec3(tmp) || !vec3(tmp)) return false;
    } else if (m == "UniformScaleTranslateMap" || m == "ScaleTranslateMap") {
        if (!vec3(t) || !vec3(scale) || !vec3(vs) || !vec3(tmp) || !vec3(tmp) || !vec3(tmp)) return false;
    } else if (m == "TranslationMap") {
        if (!vec3(t)) return false;
        vs[0] = vs[1] = vs[2] = 1.0;
    } else if (m == "AffineMap") {
        double mat[16];
        if (!c.bytes(mat, sizeof(mat))) return false;
        for (int k = 0; k < 3; ++k) {   // row-vector convention: rows 0-2 the axes, row 3 the translation
            vs[k] = std::sqrt(mat[4 * k] * mat[4 * k] + mat[4 * k + 1] * mat[4 * k + 1] + mat[4 * k + 2] * mat[4 * k + 2]);
            t[k] = mat[12 + k];
        }
    } else {
        return c.fail("unsupported transform map " + m);
    }
    std::copy(vs, vs + 3, g.voxel_size);
    std::copy(t, t + 3, g.translation);
    return true;
}

}  // namespace

bool read_vdb_buffer(const uint8_t* data, size_t size, VdbFile& out, std::string& err) {
    err.clear();
    out = VdbFile{};
    Cursor c{data, size, 0, &err};
    int64_t magic;
    if (!c.get(magic)) return false;
    if (magic != 0x56444220) return c.fail("not an OpenVDB file (bad magic number)");
    if (!c.get(out.file_version)) return false;
    if (out.file_version < 222 || out.file_version > 224)
        return c.fail("unsupported OpenVDB file version " + std::to_string(out.file_version) + " (222-224 are read)");
    uint8_t has_offsets;
    if (!c.get(out.lib_major) || !c.get(out.lib_minor) || !c.get(has_offsets)) return false;
    if (!has_offsets) return c.fail("files without grid offsets are not supported");
    if (!c.need(36)) return false;
    out.uuid.assign(reinterpret_cast<const char*>(data + c.pos), 36);
    c.pos += 36;
    if (!read_meta(c, nullptr)) return false;
    uint32_t ngrids;
    if (!c.get(ngrids)) return false;
    std::map<std::string, size_t> by_unique;
    for (uint32_t gi = 0; gi < ngrids; ++gi) {
        VdbGrid g;
        std::string unique, parent;
        int64_t gpos, bpos, epos;
        if (!c.str(unique) || !c.str(g.type)) return false;
        g.name = unique.substr(0, unique.find('\x1e'));
        const std::string half_suffix = "_HalfFloat";
        if (g.type.size() > half_suffix.size() &&
            g.type.compare(g.type.size() - half_suffix.size(), half_suffix.size(), half_suffix) == 0) {
            g.half_float = true;
            g.type.resize(g.type.size() - half_suffix.size());
        }
        if (!c.str(parent) || !c.get(gpos) || !c.get(bpos) || !c.get(epos)) return false;
        if (gpos < 0 || bpos < gpos || epos < bpos || size_t(epos) > size) return c.fail("bad grid offsets");
        g.ncomp = g.type == "Tree_float_5_4_3" ? 1 : g.type == "Tree_vec3s_5_4_3" ? 3 : 0;
        if (g.ncomp && parent.empty()) {
            Cursor gc{data, size_t(epos), size_t(gpos), &err};
            uint32_t comp;
            if (!gc.get(comp) || !read_meta(gc, &g) || !read_transform(gc, g)) return false;
            GridReader r(gc, g, comp, out.file_version);
            if (!r.read(size_t(bpos))) return false;
            g.decoded = true;
        } else if (g.ncomp && !parent.empty()) {
            // an instance shares its parent's tree; its own transform and metadata
            auto it = by_unique.find(parent);
            if (it != by_unique.end() && out.grids[it->second].decoded) {
                const VdbGrid& p = out.grids[it->second];
                Cursor gc{data, size_t(epos), size_t(gpos), &err};
                uint32_t comp;
                if (!gc.get(comp) || !read_meta(gc, &g) || !read_transform(gc, g)) return false;
                g.origins = p.origins;
                g.active = p.active;
                g.values = p.values;
                g.tiles = p.tiles;
                std::copy(p.background, p.background + 3, g.background);
                g.decoded = true;
            }
        }
        by_unique[unique] = out.grids.size();
        out.grids.push_back(std::move(g));
        c.pos = size_t(epos);
    }
    return true;
}

int64_t expand_tiles(VdbGrid& g, int32_t max_dim, int64_t max_leaves) {
    if (!g.decoded) return 0;
    int64_t add = 0;
    for (const VdbTile& t : g.tiles)
        if (t.dim <= max_dim) add += int64_t(t.dim / 8) * (t.dim / 8) * (t.dim / 8);
    if (g.nleaf() + add > max_leaves) return -1;
    if (add == 0) return 0;
    const size_t nc = size_t(g.ncomp);
    std::vector<VdbTile> kept;
    for (const VdbTile& t : g.tiles) {
        if (t.dim > max_dim) {
            kept.push_back(t);
            continue;
        }
        const int32_t n = t.dim / 8;
        for (int32_t i = 0; i < n; ++i)
            for (int32_t j = 0; j < n; ++j)
                for (int32_t k = 0; k < n; ++k) {
                    g.origins.push_back(t.origin[0] + 8 * i);
                    g.origins.push_back(t.origin[1] + 8 * j);
                    g.origins.push_back(t.origin[2] + 8 * k);
                    for (int w = 0; w < 8; ++w) g.active.push_back(t.active ? ~uint64_t(0) : 0);
                    for (int v = 0; v < 512; ++v)
                        for (size_t c = 0; c < nc; ++c) g.values.push_back(t.value[c]);
                }
    }
    g.tiles = std::move(kept);
    // tree order: root cell (4096^3, by x, then y, then z), then the index in
    // the 32^3 internal node, then in the 16^3 one
    const size_t n = size_t(g.nleaf());
    auto key = [&](size_t i) {
        const int32_t* o = &g.origins[i * 3];
        const int32_t r[3] = {o[0] >> 12, o[1] >> 12, o[2] >> 12};
        const uint32_t i5 = uint32_t(((o[0] >> 7) & 31) << 10 | ((o[1] >> 7) & 31) << 5 | ((o[2] >> 7) & 31));
        const uint32_t i4 = uint32_t(((o[0] >> 3) & 15) << 8 | ((o[1] >> 3) & 15) << 4 | ((o[2] >> 3) & 15));
        return std::make_tuple(r[0], r[1], r[2], i5, i4);
    };
    std::vector<size_t> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return key(a) < key(b); });
    std::vector<int32_t> org(n * 3);
    std::vector<uint64_t> act(n * 8);
    std::vector<float> val(n * 512 * nc);
    for (size_t i = 0; i < n; ++i) {
        const size_t s = idx[i];
        std::copy(&g.origins[s * 3], &g.origins[s * 3] + 3, &org[i * 3]);
        std::copy(&g.active[s * 8], &g.active[s * 8] + 8, &act[i * 8]);
        std::copy(&g.values[s * 512 * nc], &g.values[s * 512 * nc] + 512 * nc, &val[i * 512 * nc]);
    }
    g.origins = std::move(org);
    g.active = std::move(act);
    g.values = std::move(val);
    return add;
}

bool read_vdb_file(const std::string& path, VdbFile& out, std::string& err) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    const std::streamsize n = f.tellg();
    if (n < 0) {
        err = "cannot read " + path;
        return false;
    }
    std::vector<uint8_t> buf(static_cast<size_t>(n));
    f.seekg(0);
    if (n > 0 && !f.read(reinterpret_cast<char*>(buf.data()), n)) {
        err = "cannot read " + path;
        return false;
    }
    return read_vdb_buffer(buf.data(), buf.size(), out, err);
}

}  // namespace uaamg
