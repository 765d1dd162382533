// C ABI of the OpenVDB file reader (include/uaamg_capi.h, OpenVDB files).
#include <memory>
#include <string>

#include "capi_common.h"
#include "vdb_file.h"

using namespace uaamg;
using uaamg::capi::fail;

struct uaamg_vdb_file {
    VdbFile f;
};

extern "C" {

int uaamg_vdb_open(const char* path, uaamg_vdb_file** out) {
    UAAMG_GUARD_BEGIN
    if (!out) return fail(UAAMG_EINVAL, "null output handle");
    *out = nullptr;
    if (!path) return fail(UAAMG_EINVAL, "null path");
    std::unique_ptr<uaamg_vdb_file> h(new uaamg_vdb_file());
    std::string err;
    if (!read_vdb_file(path, h->f, err)) return fail(UAAMG_EINVAL, err);
    *out = h.release();
    return 0;
    UAAMG_GUARD_END
}

void uaamg_vdb_close(uaamg_vdb_file* h) { delete h; }

int32_t uaamg_vdb_grid_count(const uaamg_vdb_file* h) { return h ? int32_t(h->f.grids.size()) : -1; }

int uaamg_vdb_grid_info(const uaamg_vdb_file* h, int32_t index, uaamg_vdb_grid_desc* info) {
    if (!h || !info) return fail(UAAMG_EINVAL, "null argument");
    if (index < 0 || size_t(index) >= h->f.grids.size()) return fail(UAAMG_EINVAL, "grid index out of range");
    const VdbGrid& g = h->f.grids[size_t(index)];
    info->name = g.name.c_str();
    info->type = g.type.c_str();
    info->grid_class = g.grid_class.c_str();
    info->decoded = g.decoded ? 1 : 0;
    info->ncomp = g.ncomp;
    info->half_float = g.half_float ? 1 : 0;
    info->nleaf = g.nleaf();
    info->ntiles = int64_t(g.tiles.size());
    for (int k = 0; k < 3; ++k) {
        info->voxel_size[k] = g.voxel_size[k];
        info->translation[k] = g.translation[k];
        info->background[k] = g.background[k];
    }
    info->meta_voxel_count = g.meta_voxel_count;
    for (int k = 0; k < 6; ++k) info->meta_bbox[k] = g.meta_bbox[k];
    info->has_meta_bbox = g.has_meta_bbox ? 1 : 0;
    return 0;
}

int uaamg_vdb_grid(const uaamg_vdb_file* h, int32_t index, uaamg_grid* out) {
    if (!h || !out) return fail(UAAMG_EINVAL, "null argument");
    if (index < 0 || size_t(index) >= h->f.grids.size()) return fail(UAAMG_EINVAL, "grid index out of range");
    const VdbGrid& g = h->f.grids[size_t(index)];
    if (!g.decoded) return fail(UAAMG_EINVAL, "grid type " + g.type + " is not decoded");
    out->nleaf = g.nleaf();
    out->origins = g.origins.data();
    out->active = g.active.data();
    out->values = g.values.data();
    out->ncomp = g.ncomp;
    for (int k = 0; k < 3; ++k) out->background[k] = g.background[k];
    return 0;
}

int uaamg_vdb_grid_tiles(const uaamg_vdb_file* h, int32_t index, int32_t* origins, int32_t* dims, float* values,
                         int32_t* active) {
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    if (index < 0 || size_t(index) >= h->f.grids.size()) return fail(UAAMG_EINVAL, "grid index out of range");
    const VdbGrid& g = h->f.grids[size_t(index)];
    for (size_t i = 0; i < g.tiles.size(); ++i) {
        const VdbTile& t = g.tiles[i];
        if (origins)
            for (int k = 0; k < 3; ++k) origins[i * 3 + size_t(k)] = t.origin[k];
        if (dims) dims[i] = t.dim;
        if (values)
            for (int k = 0; k < 3; ++k) values[i * 3 + size_t(k)] = t.value[k];
        if (active) active[i] = t.active ? 1 : 0;
    }
    return 0;
}

int uaamg_vdb_expand_tiles(uaamg_vdb_file* h, int32_t index, int32_t max_dim, int64_t max_leaves, int64_t* added) {
    UAAMG_GUARD_BEGIN
    if (!h) return fail(UAAMG_EINVAL, "null handle");
    if (index < 0 || size_t(index) >= h->f.grids.size()) return fail(UAAMG_EINVAL, "grid index out of range");
    const int64_t n = expand_tiles(h->f.grids[size_t(index)], max_dim, max_leaves);
    if (n < 0) return fail(UAAMG_ENOMEM, "expanding the tiles would exceed max_leaves");
    if (added) *added = n;
    return 0;
    UAAMG_GUARD_END
}

int uaamg_blosc_decompress(const void* src, int64_t srclen, void* dst, int64_t dstlen) {
    if (!src || !dst || srclen < 0 || dstlen < 0) return fail(UAAMG_EINVAL, "bad argument");
    std::string err;
    if (!blosc_decompress(static_cast<const uint8_t*>(src), size_t(srclen), static_cast<uint8_t*>(dst), size_t(dstlen),
                          err))
        return fail(UAAMG_EINVAL, err);
    return 0;
}

}  // extern "C"
