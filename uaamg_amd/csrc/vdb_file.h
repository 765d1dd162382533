// Reader of OpenVDB files (.vdb) into flattened leaf grids, so scenes written
// by FLIP (FLIP_vdb/main.cpp:108-131,161-185: the source / collision level
// sets and velocity grids it reads with openvdb::io::File;
// main.cpp:266-284: the sdfvel / flip_particle files it writes) can be fed to
// the solvers without OpenVDB.
//
// Format restated from OpenVDB's published file layout (the library is not
// part of this build; file versions 222-224, the ones written by OpenVDB 6-10
// and Houdini 18-20): a header (magic, file and library versions, a
// grid-offsets flag, the file UUID, file metadata), then per grid a
// descriptor (name, type, instance parent, byte offsets) and, at the grid's
// offset, its compression flags, metadata, transform, tree topology (root
// tiles and children, each internal node's child and value masks and its
// compressed tile values, each leaf's value mask) and then the leaf buffers in
// the same order (value mask, a per-node metadata byte saying how the inactive
// values were dropped, and the compressed active values). Compression: none,
// zlib ("zip"), or Blosc 1 (LZ4 or zlib codec, byte shuffle) - the Blosc
// container and the LZ4 block format are decoded here from their published
// specifications. Grids of type Tree_float_5_4_3 and Tree_vec3s_5_4_3 (full
// or half-float) are decoded; other grid types (point data, integer or mask
// grids) are listed and skipped.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace uaamg {

struct VdbTile {
    int32_t origin[3];
    int32_t dim;          // edge in voxels (8 = one leaf, 128 or 4096)
    float value[3];
    bool active;
};

struct VdbGrid {
    std::string name;               // grid name (any uniquifying suffix removed)
    std::string type;               // tree type, e.g. "Tree_float_5_4_3"
    std::string grid_class;         // "level set", "fog volume", ... (metadata "class"; "" when absent)
    std::string map_type;           // transform map, e.g. "UniformScaleMap"
    double voxel_size[3] = {0, 0, 0};
    double translation[3] = {0, 0, 0};
    bool half_float = false;
    bool decoded = false;           // false: a grid type this reader skips
    int ncomp = 0;                  // 1 (float) or 3 (vec3s)
    float background[3] = {0, 0, 0};
    int64_t meta_voxel_count = -1;  // metadata "file_voxel_count" when present
    int32_t meta_bbox[6] = {0, 0, 0, 0, 0, 0};   // "file_bbox_min" / "file_bbox_max" when present
    bool has_meta_bbox = false;
    std::vector<int32_t> origins;   // [nleaf][3], tree order
    std::vector<uint64_t> active;   // [nleaf][8] value masks (bit 64x+8y+z)
    std::vector<float> values;      // [nleaf][512][ncomp]
    std::vector<VdbTile> tiles;     // tiles of the root and internal nodes (value != background or active)
    int64_t nleaf() const { return int64_t(origins.size() / 3); }
};

struct VdbFile {
    uint32_t file_version = 0, lib_major = 0, lib_minor = 0;
    std::string uuid;
    std::vector<VdbGrid> grids;
};

// Reads every grid of the file (grids of other types keep decoded = false).
// Returns false with err set on an unreadable or malformed file.
bool read_vdb_file(const std::string& path, VdbFile& out, std::string& err);
// The same from a file image in memory.
bool read_vdb_buffer(const uint8_t* data, size_t size, VdbFile& out, std::string& err);

// Turns the grid's tiles of edge <= max_dim (8: leaf-sized tiles only; 128
// or 4096 for larger ones) into leaves holding the tile's value (all voxels
// active when the tile is, none otherwise), so that lookups outside the
// original leaves see the tile values an OpenVDB accessor returns (a level
// set's interior tiles hold -background) instead of the background. The
// leaves stay in tree order. Returns the number of leaves added, or -1 when
// that would exceed max_leaves (nothing changed).
int64_t expand_tiles(VdbGrid& g, int32_t max_dim, int64_t max_leaves);

// Decoders, exposed for their unit tests. Each returns false on malformed
// input (never reads or writes out of bounds).
bool lz4_block_decompress(const uint8_t* src, size_t srclen, uint8_t* dst, size_t dstlen, size_t* written);
bool blosc_decompress(const uint8_t* src, size_t srclen, uint8_t* dst, size_t dstlen, std::string& err);

}  // namespace uaamg
