// Device-resident variational viscosity solver (see viscosity_solver.h).
#include "viscosity_solver.h"

#include "build_timing.h"

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <utility>

namespace uaamg {

namespace {
constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
}  // namespace

ViscDevice::~ViscDevice() { release_device(); }

void ViscDevice::release_device() {
    // frees synchronise the device: kept apart from graph captures on other threads
    std::lock_guard<std::recursive_mutex> lk(capture_lock());
    if (stream_) (void)hipStreamSynchronize(stream_);
    destroy_graphs();
    for (float* p : spai0_bufs_) (void)hipFree(p);
    spai0_bufs_.clear();
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
    own_stream_ = false;
    if (mailbox_) (void)hipHostFree(mailbox_);
    mailbox_ = nullptr;
    mailbox_dev_ = nullptr;
    nu_seq_ = 0;
    lv_.clear();
    x0_ = b0_ = r0_ = p0_ = z0_ = nullptr;
    for (auto& s : staging_) s = nullptr;
    sc_ = nullptr;
    partials_ = nullptr;
    cv_ = ViscCoarsestView{};
    arena_.release();
}

int ViscDevice::check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (first_error_ == hipSuccess) {
            first_error_ = e;
            err_ = std::string(what) + ": " + hipGetErrorString(e);
        }
        return -3;
    }
    return 0;
}

void ViscDevice::init_host_only(ViscHierarchyHost&& h, const ViscOptions& opt) {
    host_ = std::move(h);
    opt_ = opt;
}

std::vector<int32_t> visc_wave_records(const ViscLevelHost& h) {
    std::vector<int32_t> rec;
    rec.reserve((h.chan_leaves[0].size() + h.chan_leaves[1].size() + h.chan_leaves[2].size()) * kVRec);
    auto slot_of = [&](int32_t lf, int gid) { return lf < 0 ? -1 : h.slot[size_t(lf) * kVSlots + size_t(gid)]; };
    auto nbk = [](int ax, int ay, int az) { return (ax + 1) * 9 + (ay + 1) * 3 + (az + 1); };
    for (int c = 0; c < 3; ++c) {
        const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
        const int own[14] = {vneg(c, 0),         vneg(c, 1),         vneg(c, 2),         vcross(d2, 0, 0),
                             vcross(d2, 0, 1),   vcross(d2, 1, 0),   vcross(d2, 1, 1),   vcross(d1, 0, 0),
                             vcross(d1, 0, 1),   vcross(d1, 1, 0),   vcross(d1, 1, 1),   kVDiag + c,
                             kVInvDiag + c,      kVSpai0 + c};
        const int mx = c == 0 ? -1 : 0, my = c == 1 ? -1 : 0, mz = c == 2 ? -1 : 0;   // -e_c
        for (int32_t leaf : h.chan_leaves[c]) {
            int32_t r[kVRec];
            std::fill(r, r + kVRec, -1);
            const int32_t* nb = &h.nb27[size_t(leaf) * 27];
            r[kVRecLeaf] = leaf;
            for (int k = 0; k < 27; ++k) r[kVRecNb + k] = nb[k];
            r[kVRecNb + nbk(0, 0, 0)] = leaf;
            for (int i = 0; i < 14; ++i) r[kVRecOwn + i] = slot_of(leaf, own[i]);
            const int32_t nm = nb[nbk(mx, my, mz)];
            r[kVRecNbr + 0] = slot_of(nb[nbk(1, 0, 0)], vneg(c, 0));
            r[kVRecNbr + 1] = slot_of(nb[nbk(0, 1, 0)], vneg(c, 1));
            r[kVRecNbr + 2] = slot_of(nb[nbk(0, 0, 1)], vneg(c, 2));
            r[kVRecNbr + 3] = slot_of(nm, vcross(d2, 1, 0));
            r[kVRecNbr + 4] = slot_of(nm, vcross(d2, 1, 1));
            r[kVRecNbr + 5] = slot_of(nm, vcross(d1, 0, 1));
            r[kVRecNbr + 6] = slot_of(nm, vcross(d1, 1, 1));
            rec.insert(rec.end(), r, r + kVRec);
        }
    }
    return rec;
}

int ViscDevice::upload_level(const ViscLevelHost& h, DevLevel& d) {
    d.bf = bf_level(h.level);
    auto put = [&](const void* src, size_t bytes) -> void* {
        void* p = arena_.take(std::max<size_t>(bytes, 4));
        if (p && bytes) check(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, stream_), "upload");
        return p;
    };
    ViscLevelView& v = d.v;
    v.nleaf = int32_t(h.nleaf());
    v.nleaf_dof = int32_t(h.nleaf_dof);
    v.nb27 = static_cast<const int32_t*>(put(h.nb27.data(), h.nb27.size() * 4));
    v.mask = static_cast<const uint64_t*>(put(h.mask.data(), h.mask.size() * 8));
    v.slot = static_cast<const int32_t*>(put(h.slot.data(), h.slot.size() * 4));
    v.pool = static_cast<const float*>(put(h.pool.data(), h.pool.size() * 4));
    std::vector<int32_t> cl;
    v.chan_off[0] = 0;
    for (int c = 0; c < 3; ++c) {
        cl.insert(cl.end(), h.chan_leaves[c].begin(), h.chan_leaves[c].end());
        v.chan_off[c + 1] = int32_t(cl.size());
    }
    v.chan_leaves = static_cast<const int32_t*>(put(cl.data(), cl.size() * 4));
    v.wide = visc_wide(v);
    v.half = visc_half(v);
    const std::vector<int32_t> wrec = visc_wave_records(h);
    v.wrec = static_cast<const int32_t*>(put(wrec.data(), wrec.size() * 4));
    v.parent = h.parent.empty() ? nullptr : static_cast<const int32_t*>(put(h.parent.data(), h.parent.size() * 4));
    v.child = h.child.empty() ? nullptr : static_cast<const int32_t*>(put(h.child.data(), h.child.size() * 4));
    std::copy(h.dflt, h.dflt + kVSlots, v.dflt);
    // flat view: 3 * nleaf_dof bricks with their own masks and DOF bases
    LevelView& f = d.flat;
    f.nleaf = int32_t(3 * h.nleaf_dof);
    f.level = h.level;
    f.mask = v.mask;
    f.dofbase = static_cast<const int64_t*>(put(h.dofbase.data(), h.dofbase.size() * 8));
    const size_t vec = std::max<size_t>(h.nleaf_dof * 3 * 512 * (d.bf ? 2 : 4), 16);
    d.X = arena_.take(vec);
    d.B = arena_.take(vec);
    d.T = arena_.take(vec);
    if (!v.nb27 || !v.mask || !v.slot || !v.pool || !v.chan_leaves || !v.wrec || !f.dofbase || !d.X || !d.B ||
        !d.T) {
        err_ = "device arena too small";
        return -2;
    }
    check(hipMemsetAsync(d.X, 0, vec, stream_), "memset");
    check(hipMemsetAsync(d.B, 0, vec, stream_), "memset");
    check(hipMemsetAsync(d.T, 0, vec, stream_), "memset");
    return first_error_ == hipSuccess ? 0 : -3;
}

bool visc_coarsest_host(const ViscHierarchyHost& H, const int32_t* leaf_map, ViscCoarsestHost& out,
                        std::string& err) {
    const ViscLevelHost& L = H.levels.back();
    const int64_t n = L.ndof;
    if (n > kViscMaxCoarsest) {
        err = "coarsest level larger than the single-workgroup solve supports";
        return false;
    }
    out = ViscCoarsestHost{};
    out.n = n;
    out.inv.assign(size_t(n), 1.f);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = H.cptr[i]; k < H.cptr[i + 1]; ++k)
            if (H.ccol[k] == i && H.cval[k] != 0.f) out.inv[i] = 1.0f / H.cval[k];
    out.pos.assign(size_t(n), 0);
    for (size_t i = 0; i < L.nleaf_dof; ++i) {
        const int64_t li = leaf_map ? leaf_map[i] : int64_t(i);
        for (int c = 0; c < 3; ++c) {
            int64_t d = L.dofbase[i * 3 + c];
            const uint64_t* m = &L.mask[(i * 3 + c) * 8];
            for (int off = 0; off < 512; ++off)
                if (mask_test(m, off)) out.pos[size_t(d++)] = int32_t((li * 3 + c) * 512 + off);
        }
    }
    int32_t width = 0;
    for (int64_t i = 0; i < n; ++i) width = std::max<int32_t>(width, int32_t(H.cptr[i + 1] - H.cptr[i]));
    if (width > kViscMaxRow) {
        err = "coarsest row longer than the stencil allows";
        return false;
    }
    out.width = width;
    // stored at the full kViscMaxRow width, so the kernel's loads need no bound
    out.ecol.assign(static_cast<size_t>(n) * kViscMaxRow, 0);
    out.elen.assign(static_cast<size_t>(n), 0);
    out.eval.assign(size_t(n) * kViscMaxRow, 0.f);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t b = H.cptr[i], len = H.cptr[i + 1] - b;
        out.elen[size_t(i)] = int32_t(len);
        for (int32_t k = 0; k < kViscMaxRow; ++k) {
            const size_t at = size_t(k) * size_t(n) + size_t(i);
            out.ecol[at] = k < len ? H.ccol[size_t(b + k)] : int32_t(i);
            out.eval[at] = k < len ? H.cval[size_t(b + k)] : 0.f;
        }
    }
    return true;
}

int ViscDevice::build_coarsest() {
    ViscCoarsestHost ch;
    if (!visc_coarsest_host(host_, nullptr, ch, err_)) return -1;
    auto put = [&](const void* src, size_t bytes) -> void* {
        void* p = arena_.take(std::max<size_t>(bytes, 4));
        if (p && bytes) check(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, stream_), "upload");
        return p;
    };
    cv_.n = int32_t(ch.n);
    cv_.ptr = static_cast<const int64_t*>(put(host_.cptr.data(), host_.cptr.size() * 8));
    cv_.col = static_cast<const int32_t*>(put(host_.ccol.data(), host_.ccol.size() * 4));
    cv_.val = static_cast<const float*>(put(host_.cval.data(), host_.cval.size() * 4));
    cv_.invdiag = static_cast<const float*>(put(ch.inv.data(), ch.inv.size() * 4));
    cv_.ewidth = ch.width;
    cv_.ecol = static_cast<const int32_t*>(put(ch.ecol.data(), ch.ecol.size() * 4));
    cv_.eval = static_cast<const float*>(put(ch.eval.data(), ch.eval.size() * 4));
    cv_.elen = static_cast<const int32_t*>(put(ch.elen.data(), ch.elen.size() * 4));
    cv_.pos = static_cast<const int32_t*>(put(ch.pos.data(), ch.pos.size() * 4));
    cv_.max_iters = 10;
    cv_.tol = FLT_EPSILON;
    if (!cv_.ptr || !cv_.col || !cv_.val || !cv_.invdiag || !cv_.pos || !cv_.ecol || !cv_.eval || !cv_.elen) {
        err_ = "device arena too small";
        return -2;
    }
    // the host arrays die here
    if (check(hipStreamSynchronize(stream_), "coarsest upload")) return -3;
    return 0;
}

int ViscDevice::init(ViscHierarchyHost&& h, const ViscOptions& opt) {
    host_ = std::move(h);
    opt_ = opt;
    if (host_.levels.empty()) {
        err_ = "empty hierarchy";
        return -1;
    }
    return acquire_device();
}

// Device memory of the hierarchy, the PCG vectors and the coarsest system: one arena.
size_t ViscDevice::device_bytes() const {
    if (host_.levels.empty()) return 0;
    size_t bytes = 1 << 20;
    for (const auto& l : host_.levels) {
        const size_t nwave = l.chan_leaves[0].size() + l.chan_leaves[1].size() + l.chan_leaves[2].size();
        bytes += align_up(nwave * kVRec * 4);
        bytes += align_up(l.nb27.size() * 4) + align_up(l.mask.size() * 8) + align_up(l.slot.size() * 4) +
                 align_up(l.pool.size() * 4) + align_up(l.dofbase.size() * 8) + 3 * align_up(l.nleaf_dof * 4) +
                 align_up(l.parent.size() * 4) + align_up(l.child.size() * 4) + 8 * kAlign;
        bytes += 3 * align_up(std::max<size_t>(l.nleaf_dof * 3 * 512 * 4, 4));
    }
    const ViscLevelHost& l0 = host_.levels[0];
    const size_t vec0 = l0.nleaf_dof * 3 * 512 * 4;
    const size_t ndof0 = size_t(l0.ndof);
    bytes += 5 * align_up(vec0) + 3 * align_up(ndof0 * 4 + 4);
    bytes += align_up(host_.cptr.size() * 8) + align_up(host_.ccol.size() * 4) + 3 * align_up(host_.cval.size() * 4) +
             8 * kAlign;
    const size_t nc = size_t(host_.levels.back().ndof);
    bytes += 2 * align_up(nc * kViscMaxRow * 4) + align_up(nc * 4) + 2 * kAlign;   // ELL rows
    bytes += align_up(sizeof(PcgScalars)) + align_up((l0.nleaf_dof * 3 / kWavesPerBlock + 2) * 8);
    return bytes;
}

int ViscDevice::acquire_device() {
    if (on_device()) return 0;
    if (host_.levels.empty()) {
        err_ = "empty hierarchy";
        return -1;
    }
    first_error_ = hipSuccess;
    err_.clear();
    if (check(hipSetDevice(opt_.device), "hipSetDevice")) return -3;
    if (opt_.stream) stream_ = opt_.stream;
    else {
        std::lock_guard<std::recursive_mutex> lk(capture_lock());
        if (check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "stream")) return -3;
        own_stream_ = true;
    }
    use_graphs_ = !(opt_.flags & 2);
    // level-to-kernel mapping of the coarse levels: each half-visit as one
    // single-workgroup launch or as the per-pass launches. By default the
    // per-pass launches (measured faster on every level tried: at 128^3 one
    // workgroup takes 1.8 ms per level-2 half-visit pair against 0.16 ms);
    // flags bit 5 selects the single-workgroup form for levels whose widest
    // channel has at most 16 leaves, UAAMG_VSMALL_LEAVES=n at most n leaves;
    // UAAMG_VSMALL_TUNE=1 times both forms on levels of at most
    // kSmallTuneLeaves leaves and keeps the faster.
    bool tune = false;
    small_max_leaves_ = (opt_.flags & 32) ? 16 : 0;
    if (const char* env = std::getenv("UAAMG_VSMALL_LEAVES")) small_max_leaves_ = std::atoi(env);
    if (const char* env = std::getenv("UAAMG_VSMALL_TUNE")) {
        if (env[0] == '1') {
            tune = true;
            small_max_leaves_ = kSmallTuneLeaves;
        }
    }
    if (opt_.flags & 128) {
        small_max_leaves_ = 0;
        tune = false;
    }
    const int L = int(host_.levels.size());
    use_small_.assign(size_t(L), 0);
    for (int l = 1; l + 1 < L; ++l) {
        const ViscLevelHost& hl = host_.levels[size_t(l)];
        const size_t widest =
            std::max(hl.chan_leaves[0].size(), std::max(hl.chan_leaves[1].size(), hl.chan_leaves[2].size()));
        use_small_[size_t(l)] = int64_t(widest) <= small_max_leaves_ ? 1 : 0;
    }
    const ViscLevelHost& l0 = host_.levels[0];
    const size_t vec0 = l0.nleaf_dof * 3 * 512 * 4;
    const size_t ndof0 = size_t(l0.ndof);
    StageClock clock;
    if (hipError_t e = arena_.reserve(device_bytes())) {
        err_ = std::string("device allocation: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? -2 : -3;
    }
    clock.lap("viscosity device: arena");
    lv_.resize(L);
    for (int l = 0; l < L; ++l)
        if (int rc = upload_level(host_.levels[l], lv_[l])) return rc;
    clock.lap("viscosity device: level upload");
    auto take = [&](size_t b) { return static_cast<float*>(arena_.take(std::max<size_t>(b, 4))); };
    x0_ = take(vec0);
    b0_ = take(vec0);
    r0_ = take(vec0);
    p0_ = take(vec0);
    z0_ = take(vec0);
    for (auto& s : staging_) s = take(ndof0 * 4 + 4);
    sc_ = static_cast<PcgScalars*>(arena_.take(sizeof(PcgScalars)));
    partials_ = reinterpret_cast<double*>(arena_.take((l0.nleaf_dof * 3 / kWavesPerBlock + 2) * 8));
    if (!x0_ || !b0_ || !r0_ || !p0_ || !z0_ || !staging_[2] || !sc_ || !partials_) {
        err_ = "device arena too small";
        return -2;
    }
    for (float* p : {x0_, b0_, r0_, p0_, z0_}) check(hipMemsetAsync(p, 0, std::max<size_t>(vec0, 4), stream_), "memset");
    check(hipMemsetAsync(sc_, 0, sizeof(PcgScalars), stream_), "memset");
    {
        std::lock_guard<std::recursive_mutex> lk(capture_lock());
        if (check(hipHostMalloc(reinterpret_cast<void**>(&mailbox_), sizeof(HostMailbox),
                                hipHostMallocMapped | hipHostMallocCoherent),
                  "mailbox"))
            return -3;
    }
    mailbox_->word = 0;
    if (check(hipHostGetDevicePointer(reinterpret_cast<void**>(&mailbox_dev_), mailbox_, 0), "mailbox map")) return -3;
    if (int rc = build_coarsest()) return rc;
    if (check(hipStreamSynchronize(stream_), "init sync")) return -3;
    clock.lap("viscosity device: vectors + coarsest");
    if (tune) tune_levels();
    if (check(hipStreamSynchronize(stream_), "init sync")) return -3;
    clock.lap("viscosity device: level mapping timing");
    return first_error_ == hipSuccess ? 0 : -3;
}

// One half-visit pair of level l's own work (the coarser levels' work between
// them left out) in either mapping, on the level's zero scratch vectors (zero
// in, zero out).
template <class V>
void ViscDevice::half_visit(int l, bool small) {
    DevLevel& d = lv_[size_t(l)];
    DevLevel& c = lv_[size_t(l + 1)];
    V* X = static_cast<V*>(d.X);
    V* B = static_cast<V*>(d.B);
    V* T = static_cast<V*>(d.T);
    V* cX = static_cast<V*>(c.X);   // tuned levels are >= 1: the coarse level has the same storage
    V* cB = static_cast<V*>(c.B);
    if (small) {
        check(launch_visc_small_pre(d.v, c.v, X, B, T, cB, true, stream_), "tune");
        check(launch_visc_small_post(d.v, X, B, cX, stream_), "tune");
        return;
    }
    check(launch_visc_zero(d.v, X, stream_), "tune");
    for (int i = 0; i < 2; ++i) rbgs<V>(l, X, B, true);
    apply_all<V>(l, kVResidual, X, B, T);
    check(launch_visc_restrict(c.v, d.v, static_cast<const V*>(T), cB, stream_), "tune");
    check(launch_visc_prolong(d.v, c.v, static_cast<const V*>(cX), X, stream_), "tune");
    for (int i = 0; i < 2; ++i) rbgs<V>(l, X, B, false);
}

void ViscDevice::half_visit_pair(int l, bool small) {
    if (lv_[size_t(l)].bf) half_visit<bf16>(l, small);
    else half_visit<float>(l, small);
}

void ViscDevice::tune_levels() {
    const int L = int(lv_.size());
    tune_ms_[0].assign(size_t(L), -1.f);
    tune_ms_[1].assign(size_t(L), -1.f);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (check(hipEventCreate(&e0), "event") || check(hipEventCreate(&e1), "event")) return;
    constexpr int kReps = 6;
    for (int l = 1; l + 1 < L; ++l) {
        if (!use_small_[size_t(l)]) continue;
        float ms[2] = {0.f, 0.f};
        for (int variant = 0; variant < 2; ++variant) {
            half_visit_pair(l, variant == 0);
            check(hipEventRecord(e0, stream_), "event");
            for (int r = 0; r < kReps; ++r) half_visit_pair(l, variant == 0);
            check(hipEventRecord(e1, stream_), "event");
            check(hipEventSynchronize(e1), "event");
            check(hipEventElapsedTime(&ms[variant], e0, e1), "event");
            ms[variant] /= kReps;
        }
        tune_ms_[0][size_t(l)] = ms[0];
        tune_ms_[1][size_t(l)] = ms[1];
        use_small_[size_t(l)] = ms[0] < ms[1] ? 1 : 0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

int ViscDevice::upload_spai0(int level) {
    ViscLevelHost& h = host_.levels[level];
    if (h.spai0) return 0;
    build_visc_spai0(h);
    float* pool = nullptr;
    int32_t* slot = nullptr;
    if (check(hipMalloc(reinterpret_cast<void**>(&pool), std::max<size_t>(h.pool.size() * 4, 4)), "spai0 alloc"))
        return -3;
    spai0_bufs_.push_back(pool);
    if (check(hipMalloc(reinterpret_cast<void**>(&slot), std::max<size_t>(h.slot.size() * 4, 4)), "spai0 alloc"))
        return -3;
    spai0_bufs_.push_back(reinterpret_cast<float*>(slot));
    check(hipMemcpyAsync(pool, h.pool.data(), h.pool.size() * 4, hipMemcpyHostToDevice, stream_), "spai0 upload");
    check(hipMemcpyAsync(slot, h.slot.data(), h.slot.size() * 4, hipMemcpyHostToDevice, stream_), "spai0 upload");
    // the wave records carry the new SPAI0 slots
    const std::vector<int32_t> rec = visc_wave_records(h);
    int32_t* wrec = nullptr;
    if (check(hipMalloc(reinterpret_cast<void**>(&wrec), std::max<size_t>(rec.size() * 4, 4)), "spai0 alloc"))
        return -3;
    spai0_bufs_.push_back(reinterpret_cast<float*>(wrec));
    check(hipMemcpyAsync(wrec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice, stream_), "spai0 upload");
    lv_[level].v.pool = pool;
    lv_[level].v.slot = slot;
    lv_[level].v.wrec = wrec;
    return check(hipStreamSynchronize(stream_), "spai0 sync");
}

// ------------------------------------------------------------ W-cycle
template <class V>
void ViscDevice::apply_all(int level, int mode, const V* x, const V* b, V* y) {
    if (mode != kVRedGS && mode != kVBlackGS && x != y) {
        check(launch_visc_apply_all(lv_[level].v, mode, x, b, y, stream_), "apply");
        return;
    }
    for (int c = 0; c < 3; ++c) check(launch_visc_apply(lv_[level].v, mode, c, x, b, y, stream_), "apply");
}

// XYZ_RBGS_apply / ZYX_RBGS_apply (:2221-2302): six colour passes in place
template <class V>
void ViscDevice::rbgs(int level, V* x, const V* b, bool xyz) {
    const ViscLevelView& v = lv_[level].v;
    if (xyz) {
        for (int c = 0; c < 3; ++c) {
            check(launch_visc_apply(v, kVRedGS, c, x, b, x, stream_), "red");
            check(launch_visc_apply(v, kVBlackGS, c, x, b, x, stream_), "black");
        }
    } else {
        for (int c = 2; c >= 0; --c) {
            check(launch_visc_apply(v, kVBlackGS, c, x, b, x, stream_), "black");
            check(launch_visc_apply(v, kVRedGS, c, x, b, x, stream_), "red");
        }
    }
}

// simd_viscosity3d::mucycle (:3828-3884) with mu = 2, n = 2. At the level
// above the coarsest the second visit would solve the same right-hand side
// from zero again (solve_exact starts from a zero guess) and is skipped.
template <class V>
void ViscDevice::mucycle(int level, V* x, const V* b, bool precon) {
    const int L = int(lv_.size());
    if (level == L - 1) {
        check(launch_visc_coarsest(cv_, b, x, stream_), "coarsest");
        return;
    }
    const ViscLevelView& v = lv_[level].v;
    V* T = static_cast<V*>(lv_[level].T);
    if (level >= 1 && use_small_[size_t(level)]) {
        DevLevel& c = lv_[level + 1];   // below level 0 every level has the same storage
        V* cX = static_cast<V*>(c.X);
        V* cB = static_cast<V*>(c.B);
        check(launch_visc_small_pre(v, c.v, x, b, T, cB, precon, stream_), "small pre");
        mucycle<V>(level + 1, cX, cB, true);
        if (level + 1 < L - 1) mucycle<V>(level + 1, cX, cB, false);
        check(launch_visc_small_post(v, x, b, cX, stream_), "small post");
        return;
    }
    if (precon) check(launch_visc_zero(v, x, stream_), "zero");
    for (int i = 0; i < 2; ++i) rbgs<V>(level, x, b, true);
    apply_all<V>(level, kVResidual, x, b, T);
    DevLevel& c = lv_[level + 1];
    // the coarse level's storage: BF16 below level 0 in the mixed form
    auto below = [&](auto* cX, auto* cB) {
        using VC = std::remove_pointer_t<decltype(cX)>;
        check(launch_visc_restrict(c.v, v, static_cast<const V*>(T), cB, stream_), "restrict");
        mucycle<VC>(level + 1, cX, cB, true);
        if (level + 1 < L - 1) mucycle<VC>(level + 1, cX, cB, false);
        check(launch_visc_prolong(v, c.v, static_cast<const VC*>(cX), x, stream_), "prolong");
    };
    if constexpr (std::is_same_v<V, bf16>) {
        below(static_cast<bf16*>(c.X), static_cast<bf16*>(c.B));
    } else {
        if (c.bf) below(static_cast<bf16*>(c.X), static_cast<bf16*>(c.B));
        else below(static_cast<float*>(c.X), static_cast<float*>(c.B));
    }
    for (int i = 0; i < 2; ++i) rbgs<V>(level, x, b, false);
}

// ---------------------------------------------------------------- PCG
// pcg_solve (:3694-3740) in three captured pieces:
//   init   r = b - A x, ||r||_inf
//   first  p = M r, rho = <p, r>, z = A p, alpha = rho / <p, z>, r -= alpha z, ||r||_inf
//   iter   z = M r, rho_new, beta, x += alpha p, p = z + beta p,
//          z = A p, alpha, r -= alpha z, ||r||_inf
void ViscDevice::issue_init() {
    const DevLevel& d = lv_[0];
    apply_all<float>(0, kVResidual, x0_, b0_, r0_);
    check(launch_absmax(d.flat, r0_, sc_, partials_, mailbox_dev_, stream_), "absmax");
}

void ViscDevice::issue_first() {
    const DevLevel& d = lv_[0];
    mucycle<float>(0, p0_, r0_, true);
    check(launch_dot(d.flat, p0_, r0_, sc_, partials_, kDotRho, stream_), "rho");
    apply_all<float>(0, kVNormal, p0_, p0_, z0_);
    check(launch_dot(d.flat, p0_, z0_, sc_, partials_, kDotSigma, stream_), "sigma");
    check(launch_axpy_absmax(d.flat, z0_, r0_, sc_, partials_, mailbox_dev_, stream_), "axpy");
}

void ViscDevice::issue_iter() {
    const DevLevel& d = lv_[0];
    mucycle<float>(0, z0_, r0_, true);
    check(launch_dot(d.flat, z0_, r0_, sc_, partials_, kDotRhoNew, stream_), "rho_new");
    check(launch_update_xp(d.flat, z0_, x0_, p0_, sc_, stream_), "update");
    apply_all<float>(0, kVNormal, p0_, p0_, z0_);
    check(launch_dot(d.flat, p0_, z0_, sc_, partials_, kDotSigma, stream_), "sigma");
    check(launch_axpy_absmax(d.flat, z0_, r0_, sc_, partials_, mailbox_dev_, stream_), "axpy");
}

void ViscDevice::destroy_graphs() {
    for (Graph* g : {&g_init_, &g_first_, &g_iter_}) {
        if (g->e) (void)hipGraphExecDestroy(g->e);
        if (g->g) (void)hipGraphDestroy(g->g);
        g->e = nullptr;
        g->g = nullptr;
    }
    graphs_ready_ = false;
}

int ViscDevice::capture(Graph& G, const std::function<void()>& body) {
    hipGraph_t g = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lk(capture_lock());
        if (check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture")) return -3;
        body();
        if (check(hipStreamEndCapture(stream_, &g), "end capture")) return -3;
    }
    G.g = g;
    if (check(hipGraphInstantiate(&G.e, G.g, nullptr, nullptr, 0), "instantiate")) return -3;
    return 0;
}

int ViscDevice::ensure_graphs() {
    if (graphs_ready_) return 0;
    int rc = capture(g_init_, [&] { issue_init(); });
    if (!rc) rc = capture(g_first_, [&] { issue_first(); });
    if (!rc) rc = capture(g_iter_, [&] { issue_iter(); });
    if (rc) {
        destroy_graphs();
        return rc;
    }
    graphs_ready_ = true;
    return 0;
}

int ViscDevice::run(Graph& g, const std::function<void()>& body) {
    if (use_graphs_) return check(hipGraphLaunch(g.e, stream_), "graph launch");
    body();
    return first_error_ == hipSuccess ? 0 : -3;
}

int ViscDevice::wait_nu(float* nu) {
    const unsigned int want = ++nu_seq_;
    uint64_t word = 0;
    for (unsigned spin = 0;; ++spin) {
        word = mailbox_->word;
        if (uint32_t(word >> 32) == want) break;
        if ((spin & 1023u) == 1023u) {
            const hipError_t q = hipStreamQuery(stream_);
            if (q == hipSuccess) {
                word = mailbox_->word;
                if (uint32_t(word >> 32) == want) break;
                err_ = "reduction result missing from the mailbox";
                return -3;
            }
            if (q != hipErrorNotReady) return check(q, "stream query");
        }
    }
    const uint32_t bits = uint32_t(word);
    std::memcpy(nu, &bits, sizeof(float));
    return 0;
}

int ViscDevice::follow_caller() {
    if (!own_stream_) return 0;
    return check(caller_.follow(stream_), "order after the caller's stream");
}

int ViscDevice::solve_pcg_dev(float* x_dof, const float* b_dof, int* iterations) {
    if (!on_device()) return -4;
    if (first_error_ != hipSuccess) return -3;
    if (use_graphs_ && ensure_graphs()) return -3;
    if (follow_caller()) return -3;
    const DevLevel& d = lv_[0];
    if (iterations) *iterations = 0;
    check(launch_scatter_dof(d.flat, x_dof, x0_, stream_), "scatter x");
    check(launch_scatter_dof(d.flat, b_dof, b0_, stream_), "scatter b");
    float nu;
    if (run(g_init_, [&] { issue_init(); })) return -3;
    if (wait_nu(&nu)) return -3;
    const float numax = opt_.tolerance * nu;
    int it = 0;
    if (!(nu <= numax) && opt_.max_iterations > 0) {
        if (run(g_first_, [&] { issue_first(); })) return -3;
        for (;;) {
            if (wait_nu(&nu)) return -3;
            if (nu <= numax || it == opt_.max_iterations - 1) {
                check(launch_update_x(d.flat, p0_, x0_, sc_, stream_), "update x");
                break;
            }
            if (run(g_iter_, [&] { issue_iter(); })) return -3;
            ++it;
        }
    }
    if (iterations) *iterations = it;
    check(launch_gather_dof(d.flat, x0_, x_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "solve sync")) return -3;
    return first_error_ == hipSuccess ? 0 : -3;
}

// ------------------------------------------------------ unit operations
int ViscDevice::apply_dev(int level, int mode, const float* x_dof, const float* b_dof, float* y_dof) {
    if (!on_device()) return -4;
    if (level < 0 || level >= nlevels() || mode < 0 || mode > 7) return -1;
    if (lv_[level].bf) {
        err_ = "unit operations run on FP32 level vectors (precision 0)";
        return -1;
    }
    if (mode == kVSpai0Mode)
        if (int rc = upload_spai0(level)) return rc;
    DevLevel& d = lv_[level];
    float *X = d.Xf(), *B = d.Bf(), *T = d.Tf();
    const size_t vec = size_t(d.v.nleaf_dof) * 3 * 512 * 4;
    check(launch_scatter_dof(d.flat, x_dof, X, stream_), "scatter");
    if (b_dof) check(launch_scatter_dof(d.flat, b_dof, B, stream_), "scatter");
    else check(hipMemsetAsync(B, 0, std::max<size_t>(vec, 4), stream_), "memset");
    if (mode <= kVSpai0Mode) {
        check(hipMemsetAsync(T, 0, std::max<size_t>(vec, 4), stream_), "memset");
        apply_all<float>(level, mode, X, B, T);
        check(launch_gather_dof(d.flat, T, y_dof, stream_), "gather");
    } else {
        if (mode == kVRedGS || mode == kVBlackGS) {
            for (int c = 0; c < 3; ++c) check(launch_visc_apply(d.v, mode, c, X, B, X, stream_), "gs");
        } else {
            rbgs<float>(level, X, B, mode == 6);
        }
        check(launch_gather_dof(d.flat, X, y_dof, stream_), "gather");
    }
    return check(hipStreamSynchronize(stream_), "apply sync");
}

int ViscDevice::restrict_dev(int fl, const float* fine_dof, float* coarse_dof) {
    if (!on_device()) return -4;
    if (fl < 0 || fl + 1 >= nlevels() || lv_[fl + 1].bf) return -1;   // FP32 level vectors only
    DevLevel &f = lv_[fl], &c = lv_[fl + 1];
    check(launch_scatter_dof(f.flat, fine_dof, f.Tf(), stream_), "scatter");
    check(hipMemsetAsync(c.B, 0, std::max<size_t>(size_t(c.v.nleaf_dof) * 3 * 512 * 4, 4), stream_), "memset");
    check(launch_visc_restrict(c.v, f.v, f.Tf(), c.Bf(), stream_), "restrict");
    check(launch_gather_dof(c.flat, c.Bf(), coarse_dof, stream_), "gather");
    return check(hipStreamSynchronize(stream_), "restrict sync");
}

int ViscDevice::prolong_dev(int fl, const float* coarse_dof, float* fine_dof) {
    if (!on_device()) return -4;
    if (fl < 0 || fl + 1 >= nlevels() || lv_[fl + 1].bf) return -1;   // FP32 level vectors only
    DevLevel &f = lv_[fl], &c = lv_[fl + 1];
    check(launch_scatter_dof(c.flat, coarse_dof, c.Xf(), stream_), "scatter");
    check(launch_scatter_dof(f.flat, fine_dof, f.Xf(), stream_), "scatter");
    check(launch_visc_prolong(f.v, c.v, c.Xf(), f.Xf(), stream_), "prolong");
    check(launch_gather_dof(f.flat, f.Xf(), fine_dof, stream_), "gather");
    return check(hipStreamSynchronize(stream_), "prolong sync");
}

int ViscDevice::coarsest_dev(const float* b_dof, float* x_dof) {
    if (!on_device()) return -4;
    DevLevel& d = lv_.back();
    if (d.bf) return -1;   // FP32 level vectors only
    const size_t vec = std::max<size_t>(size_t(d.v.nleaf_dof) * 3 * 512 * 4, 4);
    check(launch_scatter_dof(d.flat, b_dof, d.Bf(), stream_), "scatter");
    check(hipMemsetAsync(d.X, 0, vec, stream_), "memset");
    check(launch_visc_coarsest(cv_, d.Bf(), d.Xf(), stream_), "coarsest");
    check(launch_gather_dof(d.flat, d.Xf(), x_dof, stream_), "gather");
    return check(hipStreamSynchronize(stream_), "coarsest sync");
}

int ViscDevice::precondition_dev(const float* r_dof, float* z_dof) {
    if (!on_device()) return -4;
    const DevLevel& d = lv_[0];
    check(launch_scatter_dof(d.flat, r_dof, r0_, stream_), "scatter");
    mucycle<float>(0, z0_, r0_, true);
    check(launch_gather_dof(d.flat, z0_, z_dof, stream_), "gather");
    return check(hipStreamSynchronize(stream_), "precondition sync");
}

// ----------------------------------------------------------- measurement
double ViscDevice::algorithmic_bytes(int level, int mode, int c) const {
    if (level < 0 || level >= nlevels() || c < 0 || c > 2) return 0.0;
    const ViscLevelHost& L = host_.levels[level];
    const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
    std::vector<int> grids = {vneg(c, 0), vneg(c, 1), vneg(c, 2)};
    for (int t = 0; t < 4; ++t) {
        grids.push_back(vcross(d2, t >> 1, t & 1));
        grids.push_back(vcross(d1, t >> 1, t & 1));
    }
    if (mode == kVNormal || mode == kVResidual || mode == kVJacobi || mode == kVSpai0Mode) grids.push_back(kVDiag + c);
    if (mode == kVJacobi || mode == kVRedGS || mode == kVBlackGS) grids.push_back(kVInvDiag + c);
    if (mode == kVSpai0Mode) grids.push_back(kVSpai0 + c);
    const int vec_streams = 2 + (mode == kVNormal ? 0 : 1);   // x_c, y (, b)
    double bytes = 0;
    for (int32_t i : L.chan_leaves[c]) {
        const int nc = mask_count(&L.mask[(size_t(i) * 3 + c) * 8]);
        const int n1 = mask_count(&L.mask[(size_t(i) * 3 + d1) * 8]);
        const int n2 = mask_count(&L.mask[(size_t(i) * 3 + d2) * 8]);
        bytes += 4.0 * (vec_streams * nc + n1 + n2);
        for (int g : grids)
            if (L.slot[size_t(i) * kVSlots + g] >= 0) bytes += 4.0 * nc;
    }
    return bytes;
}

int ViscDevice::time_apply(int level, int mode, int c, int reps, float* ms) {
    if (!on_device()) return -4;
    if (level < 0 || level >= nlevels() || mode < 0 || mode > 5 || c < 0 || c > 2 || reps <= 0) return -1;
    DevLevel& d = lv_[level];
    if (d.bf) return -1;   // timed on FP32 level vectors
    const bool gs = mode == kVRedGS || mode == kVBlackGS;
    float* y = gs ? d.Xf() : d.Tf();
    check(launch_visc_apply(d.v, mode, c, d.Xf(), d.Bf(), y, stream_), "warm");
    hipEvent_t e0, e1;
    if (check(hipEventCreate(&e0), "event") || check(hipEventCreate(&e1), "event")) return -3;
    check(hipEventRecord(e0, stream_), "record");
    for (int r = 0; r < reps; ++r) check(launch_visc_apply(d.v, mode, c, d.Xf(), d.Bf(), y, stream_), "timed");
    check(hipEventRecord(e1, stream_), "record");
    check(hipEventSynchronize(e1), "event sync");
    float t = 0.f;
    check(hipEventElapsedTime(&t, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / float(reps);
    return first_error_ == hipSuccess ? 0 : -3;
}

}  // namespace uaamg
