// Batched variational-viscosity solve (see visc_batch.h).
#include "visc_batch.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>

namespace uaamg {

ViscBatch::~ViscBatch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    destroy_graphs();
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
    if (mailbox_) (void)hipHostFree(mailbox_);
}

void ViscBatch::destroy_graphs() {
    for (Graph* g : {&g_init_, &g_first_, &g_iter_, &g_final_}) {
        if (g->e) (void)hipGraphExecDestroy(g->e);
        if (g->g) (void)hipGraphDestroy(g->g);
        g->e = nullptr;
        g->g = nullptr;
    }
    graphs_ready_ = false;
}

int ViscBatch::check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (first_error_ == hipSuccess) {
            first_error_ = e;
            err_ = std::string(what) + ": " + hipGetErrorString(e);
        }
        return -3;
    }
    return 0;
}

int64_t ViscBatch::ndof(int input_index) const {
    if (input_index < 0 || input_index >= nproblems()) return -1;
    return ndof_[size_t(rank_[size_t(input_index)])];
}

namespace {

// the grids a channel-c wave record's slots refer to (visc_wave_records)
void record_grids(int c, int own[14], int nbr[7]) {
    const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
    const int o[14] = {vneg(c, 0),       vneg(c, 1),       vneg(c, 2),       vcross(d2, 0, 0), vcross(d2, 0, 1),
                       vcross(d2, 1, 0), vcross(d2, 1, 1), vcross(d1, 0, 0), vcross(d1, 0, 1), vcross(d1, 1, 0),
                       vcross(d1, 1, 1), kVDiag + c,       kVInvDiag + c,    kVSpai0 + c};
    const int n[7] = {vneg(c, 0), vneg(c, 1), vneg(c, 2), vcross(d2, 1, 0), vcross(d2, 1, 1), vcross(d1, 0, 1),
                      vcross(d1, 1, 1)};
    std::copy(o, o + 14, own);
    std::copy(n, n + 7, nbr);
}

}  // namespace

int ViscBatch::init(std::vector<ViscHierarchyHost>&& probs, const ViscOptions& opt) {
    opt_ = opt;
    const int K = int(probs.size());
    if (K == 0) {
        err_ = "empty batch";
        return -1;
    }
    if (opt_.precision != 0) {
        err_ = "the merged viscosity batch runs the FP32 form";
        return -1;
    }
    use_graphs_ = !(opt_.flags & 2);
    for (const auto& p : probs) {
        if (p.levels.empty() || p.levels[0].ndof == 0) {
            err_ = "a problem has no degrees of freedom";
            return -1;
        }
        if (p.levels.back().ndof > kViscMaxCoarsest) {
            err_ = "a problem's coarsest level exceeds the single-workgroup solver size";
            return -1;
        }
    }
    // merged order: deepest hierarchy first (stable in the input order)
    order_.resize(size_t(K));
    std::iota(order_.begin(), order_.end(), 0);
    std::stable_sort(order_.begin(), order_.end(),
                     [&](int a, int b) { return probs[size_t(a)].levels.size() > probs[size_t(b)].levels.size(); });
    rank_.assign(size_t(K), 0);
    nlev_.resize(size_t(K));
    ndof_.resize(size_t(K));
    for (int k = 0; k < K; ++k) {
        rank_[size_t(order_[size_t(k)])] = k;
        nlev_[size_t(k)] = int(probs[size_t(order_[size_t(k)])].levels.size());
        ndof_[size_t(k)] = probs[size_t(order_[size_t(k)])].levels[0].ndof;
    }
    const int L = nlev_[0];
    auto hl = [&](int k, int l) -> const ViscLevelHost& {
        return probs[size_t(order_[size_t(k)])].levels[size_t(l)];
    };

    // leaf offsets per level: DOF leaves of every problem (level 0 aligned to
    // the reduction block), then the coefficient-only leaves
    std::vector<std::vector<int64_t>> doff(static_cast<size_t>(L)), xoff(static_cast<size_t>(L));
    std::vector<int64_t> ndl(static_cast<size_t>(L), 0), nsm(static_cast<size_t>(L), 0), nleaf(static_cast<size_t>(L), 0);
    lv_.assign(size_t(L), Level{});
    for (int l = 0; l < L; ++l) {
        int64_t o = 0, x = 0;
        for (int k = 0; k < K && nlev_[size_t(k)] > l; ++k) {
            const ViscLevelHost& h = hl(k, l);
            doff[size_t(l)].push_back(o);
            xoff[size_t(l)].push_back(x);
            o += int64_t(h.nleaf_dof);
            if (l == 0) o = (o + kWavesPerBlock - 1) / kWavesPerBlock * kWavesPerBlock;
            x += int64_t(h.nleaf() - h.nleaf_dof);
            if (nlev_[size_t(k)] > l + 1) nsm[size_t(l)] = o;
        }
        ndl[size_t(l)] = o;
        nleaf[size_t(l)] = o + x;
        int first = -1, count = 0, maxn = 0;
        for (int k = 0; k < K; ++k)
            if (nlev_[size_t(k)] == l + 1) {
                if (first < 0) first = k;
                ++count;
                maxn = std::max<int>(maxn, int(hl(k, l).ndof));
            }
        lv_[size_t(l)].coarse_first = std::max(first, 0);
        lv_[size_t(l)].coarse_count = count;
        lv_[size_t(l)].coarse_max_n = maxn;
    }
    if (nleaf[0] * 3 > int64_t(INT32_MAX) / 512) {
        err_ = "batch too large for 32-bit brick positions";
        return -1;
    }
    auto map = [&](int k, int l, int64_t j) -> int32_t {
        if (j < 0) return -1;
        const ViscLevelHost& h = hl(k, l);
        if (size_t(j) < h.nleaf_dof) return int32_t(doff[size_t(l)][size_t(k)] + j);
        return int32_t(ndl[size_t(l)] + xoff[size_t(l)][size_t(k)] + (j - int64_t(h.nleaf_dof)));
    };

    // host images of the merged tables
    struct HostLevel {
        std::vector<uint64_t> mask;
        std::vector<int64_t> dofbase;
        std::vector<int32_t> nb27, parent, child, chan[2], rec[2];
        int32_t chan_off[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};   // [full / smoothing]
        BrickVec pool;
    };
    std::vector<HostLevel> H(static_cast<size_t>(L));
    for (int l = 0; l < L; ++l) {
        HostLevel& hm = H[size_t(l)];
        const size_t D = size_t(ndl[size_t(l)]), N = size_t(nleaf[size_t(l)]);
        hm.mask.assign(D * 24, 0);
        hm.dofbase.assign(D * 3, -1);
        hm.nb27.assign(N * 27, -1);
        hm.parent.assign(std::max<size_t>(D, 1), -1);
        hm.child.assign(std::max<size_t>(D * 8, 1), -1);
        std::vector<int32_t> pool_off, dflt_off;
        size_t bricks = 0;
        for (int k = 0; k < K && nlev_[size_t(k)] > l; ++k)
            bricks += hl(k, l).pool.size() / kLeafSize + size_t(kVSlots);
        hm.pool.resize(bricks * kLeafSize);
        size_t at = 0;
        for (int k = 0; k < K && nlev_[size_t(k)] > l; ++k) {
            const ViscLevelHost& h = hl(k, l);
            const size_t o = size_t(doff[size_t(l)][size_t(k)]);
            std::copy(h.mask.begin(), h.mask.end(), hm.mask.begin() + long(o * 24));
            std::copy(h.dofbase.begin(), h.dofbase.end(), hm.dofbase.begin() + long(o * 3));
            for (size_t j = 0; j < h.nleaf(); ++j)
                for (int q = 0; q < 27; ++q)
                    hm.nb27[size_t(map(k, l, int64_t(j))) * 27 + size_t(q)] = map(k, l, h.nb27[j * 27 + size_t(q)]);
            if (nlev_[size_t(k)] > l + 1)
                for (size_t i = 0; i < h.nleaf_dof; ++i) {
                    const int32_t p = h.parent[i];
                    if (p >= 0) hm.parent[o + i] = map(k, l + 1, p >> 3) * 8 + (p & 7);
                }
            if (l >= 1)
                for (size_t i = 0; i < h.nleaf_dof; ++i)
                    for (int b = 0; b < 8; ++b) hm.child[(o + i) * 8 + size_t(b)] = map(k, l - 1, h.child[i * 8 + size_t(b)]);
            // bricks: the problem's pool, then one brick per grid holding its default
            pool_off.push_back(int32_t(at / kLeafSize));
            std::copy(h.pool.begin(), h.pool.end(), hm.pool.begin() + long(at));
            at += h.pool.size();
            dflt_off.push_back(int32_t(at / kLeafSize));
            for (int g = 0; g < kVSlots; ++g) {
                std::fill(hm.pool.begin() + long(at), hm.pool.begin() + long(at + kLeafSize), h.dflt[g]);
                at += kLeafSize;
            }
        }
        // wave records and channel lists, every problem (0) and the smoothing prefix (1)
        std::vector<std::vector<int32_t>> recs(static_cast<size_t>(K));
        for (int k = 0; k < K && nlev_[size_t(k)] > l; ++k) recs[size_t(k)] = visc_wave_records(hl(k, l));
        for (int view = 0; view < 2; ++view) {
            for (int c = 0; c < 3; ++c) {
                int own[14], nbr[7];
                record_grids(c, own, nbr);
                for (int k = 0; k < K && nlev_[size_t(k)] > l; ++k) {
                    if (view == 1 && nlev_[size_t(k)] <= l + 1) continue;
                    const ViscLevelHost& h = hl(k, l);
                    size_t first = 0;
                    for (int cc = 0; cc < c; ++cc) first += h.chan_leaves[cc].size();
                    for (size_t e = 0; e < h.chan_leaves[c].size(); ++e) {
                        const int32_t* r = &recs[size_t(k)][(first + e) * kVRec];
                        int32_t m[kVRec];
                        std::copy(r, r + kVRec, m);
                        m[kVRecLeaf] = map(k, l, r[kVRecLeaf]);
                        for (int q = 0; q < 27; ++q) m[kVRecNb + q] = map(k, l, r[kVRecNb + q]);
                        for (int q = 0; q < 14; ++q)
                            m[kVRecOwn + q] = r[kVRecOwn + q] >= 0 ? r[kVRecOwn + q] + pool_off[size_t(k)]
                                                                   : dflt_off[size_t(k)] + own[q];
                        for (int q = 0; q < 7; ++q)
                            m[kVRecNbr + q] = r[kVRecNbr + q] >= 0 ? r[kVRecNbr + q] + pool_off[size_t(k)]
                                                                   : dflt_off[size_t(k)] + nbr[q];
                        hm.rec[view].insert(hm.rec[view].end(), m, m + kVRec);
                        hm.chan[view].push_back(m[kVRecLeaf]);
                    }
                }
                hm.chan_off[view][c + 1] = int32_t(hm.chan[view].size());
            }
        }
    }
    // coarsest systems, merged order, positions in the merged level vector
    std::vector<ViscCoarsestHost> ch(static_cast<size_t>(K));
    for (int k = 0; k < K; ++k) {
        const int c = nlev_[size_t(k)] - 1;
        const ViscLevelHost& h = hl(k, c);
        std::vector<int32_t> lm(std::max<size_t>(h.nleaf_dof, 1));
        for (size_t i = 0; i < h.nleaf_dof; ++i) lm[i] = map(k, c, int64_t(i));
        if (!visc_coarsest_host(probs[size_t(order_[size_t(k)])], lm.data(), ch[size_t(k)], err_)) return -1;
    }

    // one arena for everything
    constexpr size_t kAl = 256;
    auto al = [](size_t b) { return (std::max<size_t>(b, 4) + kAl - 1) / kAl * kAl; };
    size_t bytes = 1 << 16;
    for (int l = 0; l < L; ++l) {
        const HostLevel& h = H[size_t(l)];
        const size_t D = size_t(ndl[size_t(l)]);
        bytes += al(h.mask.size() * 8) + al(h.dofbase.size() * 8) + al(h.nb27.size() * 4) + al(h.parent.size() * 4) +
                 al(h.child.size() * 4) + al(h.pool.size() * 4);
        for (int v = 0; v < 2; ++v) bytes += al(h.rec[v].size() * 4) + al(h.chan[v].size() * 4);
        bytes += 3 * al(D * 3 * kLeafSize * 4);
    }
    const size_t D0 = size_t(ndl[0]);
    bytes += 5 * al(D0 * 3 * kLeafSize * 4);                       // x b r p z
    bytes += al(D0 * 3 * 4) + al((size_t(K) + 1) * 4);              // brick -> problem, block ranges
    bytes += al((D0 * 3 / kWavesPerBlock + 16) * 8);                // partials
    bytes += al(size_t(K) * sizeof(BatchPcg)) + al(64) + al(size_t(K) * sizeof(ViscCoarsestView));
    for (const auto& c : ch)
        bytes += al(c.inv.size() * 4) + al(c.pos.size() * 4) + al(c.ecol.size() * 4) + al(c.eval.size() * 4) +
                 al(c.elen.size() * 4);

    if (check(hipSetDevice(opt_.device), "hipSetDevice")) return -3;
    if (opt_.stream) {
        stream_ = opt_.stream;
    } else {
        if (check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate")) return -3;
        own_stream_ = true;
    }
    if (check(arena_.reserve(bytes), "hipMalloc")) return -2;
    auto up = [&](const void* src, size_t b) -> void* {
        void* p = arena_.take(std::max<size_t>(b, 4));
        if (p && b) (void)check(hipMemcpyAsync(p, src, b, hipMemcpyHostToDevice, stream_), "upload");
        return p;
    };
    auto zeros = [&](size_t b) -> void* {
        void* p = arena_.take(std::max<size_t>(b, 4));
        if (p) (void)check(hipMemsetAsync(p, 0, std::max<size_t>(b, 4), stream_), "memset");
        return p;
    };
    for (int l = 0; l < L; ++l) {
        HostLevel& h = H[size_t(l)];
        Level& d = lv_[size_t(l)];
        const size_t D = size_t(ndl[size_t(l)]);
        ViscLevelView& v = d.v;
        v.nleaf = int32_t(nleaf[size_t(l)]);
        v.nleaf_dof = int32_t(D);
        v.nb27 = static_cast<const int32_t*>(up(h.nb27.data(), h.nb27.size() * 4));
        v.mask = static_cast<const uint64_t*>(up(h.mask.data(), h.mask.size() * 8));
        v.slot = nullptr;   // the wave records carry every slot the kernels read
        v.pool = static_cast<const float*>(up(h.pool.data(), h.pool.size() * 4));
        v.parent = static_cast<const int32_t*>(up(h.parent.data(), h.parent.size() * 4));
        v.child = l >= 1 ? static_cast<const int32_t*>(up(h.child.data(), h.child.size() * 4)) : nullptr;
        // read only through slots that are never -1 here; kept for completeness
        std::copy(hl(0, l).dflt, hl(0, l).dflt + kVSlots, v.dflt);
        const int64_t* dofbase = static_cast<const int64_t*>(up(h.dofbase.data(), h.dofbase.size() * 8));
        ViscLevelView sv = v;
        for (int view = 0; view < 2; ++view) {
            ViscLevelView& t = view == 0 ? v : sv;
            t.chan_leaves = static_cast<const int32_t*>(up(h.chan[view].data(), h.chan[view].size() * 4));
            t.wrec = static_cast<const int32_t*>(up(h.rec[view].data(), h.rec[view].size() * 4));
            std::copy(h.chan_off[view], h.chan_off[view] + 4, t.chan_off);
            t.wide = visc_wide(t);
            t.half = visc_half(t);
        }
        // the smoothing prefix's DOF leaves come first: its vectors are the prefix of the level's
        sv.nleaf_dof = int32_t(nsm[size_t(l)]);
        d.sv = sv;
        LevelView& f = d.flat;
        f.nleaf = int32_t(3 * D);
        f.level = l;
        f.mask = v.mask;
        f.dofbase = dofbase;
        d.X = static_cast<float*>(zeros(D * 3 * kLeafSize * 4));
        d.B = static_cast<float*>(zeros(D * 3 * kLeafSize * 4));
        d.T = static_cast<float*>(zeros(D * 3 * kLeafSize * 4));
        if (!v.nb27 || !v.mask || !v.pool || !v.parent || !dofbase || !d.X || !d.B || !d.T || !v.wrec || !sv.wrec) {
            err_ = "arena too small";
            return -2;
        }
        // the host images die with this function: finish the copies first
        if (check(hipStreamSynchronize(stream_), "upload sync")) return -3;
    }
    x0_ = static_cast<float*>(zeros(D0 * 3 * kLeafSize * 4));
    b0_ = static_cast<float*>(zeros(D0 * 3 * kLeafSize * 4));
    r0_ = static_cast<float*>(zeros(D0 * 3 * kLeafSize * 4));
    p0_ = static_cast<float*>(zeros(D0 * 3 * kLeafSize * 4));
    z0_ = static_cast<float*>(zeros(D0 * 3 * kLeafSize * 4));
    {
        std::vector<int32_t> bp(D0 * 3, 0), blk(size_t(K) + 1, 0);
        off0_.assign(size_t(K), 0);
        sub0_.assign(size_t(K), LevelView{});
        for (int k = 0; k < K; ++k) {
            const int64_t o = doff[0][size_t(k)];
            const int64_t e = k + 1 < K ? doff[0][size_t(k + 1)] : int64_t(D0);
            for (int64_t j = 3 * o; j < 3 * e; ++j) bp[size_t(j)] = k;
            blk[size_t(k)] = int32_t(3 * o / kWavesPerBlock);
            off0_[size_t(k)] = 3 * o * kLeafSize;
            LevelView s = lv_[0].flat;
            s.nleaf = int32_t(3 * hl(k, 0).nleaf_dof);
            s.mask = lv_[0].flat.mask + 3 * o * 8;
            s.dofbase = lv_[0].flat.dofbase + 3 * o;
            sub0_[size_t(k)] = s;
        }
        blk[size_t(K)] = int32_t(3 * int64_t(D0) / kWavesPerBlock);
        brickprob_ = static_cast<const int32_t*>(up(bp.data(), bp.size() * 4));
        blk_begin_ = static_cast<const int32_t*>(up(blk.data(), blk.size() * 4));
        partials_ = static_cast<double*>(zeros((D0 * 3 / kWavesPerBlock + 16) * 8));
        st_ = static_cast<BatchPcg*>(zeros(size_t(K) * sizeof(BatchPcg)));
        seq_dev_ = static_cast<unsigned int*>(zeros(64));
        std::vector<ViscCoarsestView> views(static_cast<size_t>(K));
        for (int k = 0; k < K; ++k) {
            const ViscCoarsestHost& c = ch[size_t(k)];
            ViscCoarsestView& cv = views[size_t(k)];
            cv.n = int32_t(c.n);
            cv.invdiag = static_cast<const float*>(up(c.inv.data(), c.inv.size() * 4));
            cv.pos = static_cast<const int32_t*>(up(c.pos.data(), c.pos.size() * 4));
            cv.ecol = static_cast<const int32_t*>(up(c.ecol.data(), c.ecol.size() * 4));
            cv.eval = static_cast<const float*>(up(c.eval.data(), c.eval.size() * 4));
            cv.elen = static_cast<const int32_t*>(up(c.elen.data(), c.elen.size() * 4));
            cv.ewidth = c.width;
            cv.max_iters = 10;
            cv.tol = FLT_EPSILON;
        }
        views_ = static_cast<ViscCoarsestView*>(up(views.data(), views.size() * sizeof(ViscCoarsestView)));
        if (check(hipStreamSynchronize(stream_), "upload sync")) return -3;
    }
    if (!z0_ || !views_ || !st_ || !seq_dev_ || !partials_ || !brickprob_ || !blk_begin_) {
        err_ = "arena too small";
        return -2;
    }
    if (check(hipHostMalloc(reinterpret_cast<void**>(&mailbox_), sizeof(HostMailbox),
                            hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc mailbox"))
        return -3;
    std::memset(const_cast<HostMailbox*>(mailbox_), 0, sizeof(HostMailbox));
    if (check(hipHostGetDevicePointer(reinterpret_cast<void**>(&mailbox_dev_), mailbox_, 0), "mailbox alias"))
        return -3;
    st_host_.assign(size_t(K), BatchPcg{});
    return check(hipStreamSynchronize(stream_), "init sync");
}

// XYZ_RBGS_apply / ZYX_RBGS_apply (:2221-2302) on the smoothing problems of level l
void ViscBatch::rbgs(int l, float* x, const float* b, bool xyz) {
    const ViscLevelView& v = lv_[size_t(l)].sv;
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

// simd_viscosity3d::mucycle (:3828-3884), mu = 2, n = 2, on every problem of
// the batch: the problems whose coarsest level is l solve it on the first
// visit (the second visit's identical zero-start solve is skipped, as in the
// single-problem cycle); the others are zeroed (first visit), smoothed on the
// prefix view, recurse twice and are smoothed again. Level 0 runs on the PCG
// vectors X, B; the coarser levels on their own.
void ViscBatch::mucycle(int l, float* X, const float* B, bool first) {
    Level& d = lv_[size_t(l)];
    if (first && d.coarse_count)
        check(launch_visc_bcoarsest(views_ + d.coarse_first, d.coarse_count, d.coarse_max_n, B, X, stream_),
              "coarsest");
    if (d.sv.chan_off[3] == 0) return;   // no problem smooths here
    if (first) check(launch_visc_zero(d.sv, X, stream_), "zero");
    for (int i = 0; i < 2; ++i) rbgs(l, X, B, true);
    check(launch_visc_apply_all(d.sv, kVResidual, X, B, d.T, stream_), "residual");
    Level& c = lv_[size_t(l + 1)];
    check(launch_visc_restrict(c.v, d.v, static_cast<const float*>(d.T), c.B, stream_), "restrict");
    mucycle(l + 1, c.X, c.B, true);
    if (c.sv.chan_off[3] > 0) mucycle(l + 1, c.X, c.B, false);
    check(launch_visc_prolong(d.v, c.v, static_cast<const float*>(c.X), X, stream_), "prolong");
    for (int i = 0; i < 2; ++i) rbgs(l, X, B, false);
}

// pcg_solve (:3694-3740) per problem, in the pieces of ViscDevice: init
// r = b - A x, ||r||_inf; first p = M r, rho, z = A p, alpha, r -= alpha z,
// ||r||_inf; iter (after the x update of the problems that converged) z = M r,
// rho_new, beta, x += alpha p, p = z + beta p, z = A p, alpha, r -= alpha z,
// ||r||_inf. Each piece ends with the count of problems still iterating.
void ViscBatch::issue_init() {
    const Level& d = lv_[0];
    const int K = nproblems();
    check(launch_visc_apply_all(d.v, kVResidual, x0_, b0_, r0_, stream_), "residual");
    check(launch_absmax_partials(d.flat, r0_, partials_, stream_), "absmax");
    check(launch_bfinalize(kBFinNu0, K, blk_begin_, partials_, st_, stream_), "nu0");
    check(launch_bstatus(K, st_, seq_dev_, mailbox_dev_, stream_), "status");
}

void ViscBatch::issue_first() {
    const Level& d = lv_[0];
    const int K = nproblems();
    mucycle(0, p0_, r0_, true);
    check(launch_dot_partials(d.flat, p0_, r0_, partials_, stream_), "rho");
    check(launch_bfinalize(kBFinRho, K, blk_begin_, partials_, st_, stream_), "rho");
    check(launch_visc_apply_all(d.v, kVNormal, p0_, p0_, z0_, stream_), "A p");
    check(launch_dot_partials(d.flat, p0_, z0_, partials_, stream_), "sigma");
    check(launch_bfinalize(kBFinSigma, K, blk_begin_, partials_, st_, stream_), "sigma");
    check(launch_baxpy_absmax(d.flat, brickprob_, st_, z0_, x0_, p0_, b0_, r0_, partials_, stream_), "axpy");
    check(launch_bfinalize(kBFinNuViscFirst, K, blk_begin_, partials_, st_, stream_), "nu");
    check(launch_bstatus(K, st_, seq_dev_, mailbox_dev_, stream_), "status");
}

void ViscBatch::issue_iter() {
    const Level& d = lv_[0];
    const int K = nproblems();
    check(launch_bupdate_x(d.flat, brickprob_, st_, p0_, x0_, stream_), "update x");
    mucycle(0, z0_, r0_, true);
    check(launch_dot_partials(d.flat, z0_, r0_, partials_, stream_), "rho_new");
    check(launch_bfinalize(kBFinRhoNew, K, blk_begin_, partials_, st_, stream_), "rho_new");
    check(launch_bupdate_xp(d.flat, brickprob_, st_, z0_, x0_, p0_, stream_), "update");
    check(launch_visc_apply_all(d.v, kVNormal, p0_, p0_, z0_, stream_), "A p");
    check(launch_dot_partials(d.flat, p0_, z0_, partials_, stream_), "sigma");
    check(launch_bfinalize(kBFinSigma, K, blk_begin_, partials_, st_, stream_), "sigma");
    check(launch_baxpy_absmax(d.flat, brickprob_, st_, z0_, x0_, p0_, b0_, r0_, partials_, stream_), "axpy");
    check(launch_bfinalize(kBFinNuViscIter, K, blk_begin_, partials_, st_, stream_), "nu");
    check(launch_bstatus(K, st_, seq_dev_, mailbox_dev_, stream_), "status");
}

// the problems that converged in the last piece: x += alpha p, then commit
void ViscBatch::issue_final() {
    const Level& d = lv_[0];
    check(launch_bupdate_x(d.flat, brickprob_, st_, p0_, x0_, stream_), "update x");
    check(launch_bfinalize(-1, nproblems(), blk_begin_, partials_, st_, stream_), "commit");
}

int ViscBatch::capture(Graph& G, const std::function<void()>& body) {
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

int ViscBatch::run(Graph& g, const std::function<void()>& body) {
    if (use_graphs_) return check(hipGraphLaunch(g.e, stream_), "graph launch");
    body();
    return first_error_ == hipSuccess ? 0 : -3;
}

int ViscBatch::wait_active(int* active) {
    const unsigned int want = ++seq_;
    uint64_t word = 0;
    for (unsigned spin = 0;; ++spin) {
        word = mailbox_->word;
        if (uint32_t(word >> 32) == want) break;
        if ((spin & 1023u) == 1023u) {
            const hipError_t q = hipStreamQuery(stream_);
            if (q == hipSuccess) {
                word = mailbox_->word;
                if (uint32_t(word >> 32) == want) break;
                err_ = "batch status missing from the mailbox";
                return -3;
            }
            if (q != hipErrorNotReady) return check(q, "stream query");
        }
    }
    *active = int(uint32_t(word));
    return 0;
}

int ViscBatch::solve(float tolerance, float* const* x, const float* const* b, int* iterations) {
    if (lv_.empty()) return -4;
    if (first_error_ != hipSuccess) return -3;
    const int K = nproblems();
    for (int i = 0; i < K; ++i)
        if (!x[i] || !b[i]) {
            err_ = "null vector in the batch";
            return -1;
        }
    if (use_graphs_ && !graphs_ready_) {
        int rc = capture(g_init_, [&] { issue_init(); });
        if (!rc) rc = capture(g_first_, [&] { issue_first(); });
        if (!rc) rc = capture(g_iter_, [&] { issue_iter(); });
        if (!rc) rc = capture(g_final_, [&] { issue_final(); });
        if (rc) {
            destroy_graphs();
            return rc;
        }
        graphs_ready_ = true;
    }
    if (own_stream_ && check(caller_.follow(stream_), "order after the caller's stream")) return -3;
    for (int k = 0; k < K; ++k) {
        BatchPcg& s = st_host_[size_t(k)];
        s = BatchPcg{};
        s.tol = tolerance;
        s.max_iters = opt_.max_iterations;
        s.true_nu = -1.f;
        s.replacements = -1;
    }
    check(hipMemcpyAsync(st_, st_host_.data(), size_t(K) * sizeof(BatchPcg), hipMemcpyHostToDevice, stream_),
          "upload state");
    for (int k = 0; k < K; ++k) {
        const int i = order_[size_t(k)];
        const size_t o = size_t(off0_[size_t(k)]);
        check(launch_scatter_dof(sub0_[size_t(k)], x[i], x0_ + o, stream_), "scatter x");
        check(launch_scatter_dof(sub0_[size_t(k)], b[i], b0_ + o, stream_), "scatter b");
    }
    int active = 0;
    if (run(g_init_, [&] { issue_init(); })) return -3;
    if (wait_active(&active)) return -3;
    if (active > 0) {
        if (run(g_first_, [&] { issue_first(); })) return -3;
        // every piece is followed by its status wait (the sequence numbers stay in step)
        for (int pieces = 0;; ++pieces) {
            if (wait_active(&active)) return -3;
            if (active == 0) break;
            if (pieces > opt_.max_iterations) {   // the per-problem rules end every problem by then
                err_ = "batched PCG did not terminate";
                return -3;
            }
            if (run(g_iter_, [&] { issue_iter(); })) return -3;
        }
    }
    if (run(g_final_, [&] { issue_final(); })) return -3;
    for (int k = 0; k < K; ++k) {
        const int i = order_[size_t(k)];
        const size_t o = size_t(off0_[size_t(k)]);
        check(launch_gather_dof(sub0_[size_t(k)], x0_ + o, x[i], stream_), "gather");
    }
    check(hipMemcpyAsync(st_host_.data(), st_, size_t(K) * sizeof(BatchPcg), hipMemcpyDeviceToHost, stream_),
          "download state");
    if (check(hipStreamSynchronize(stream_), "solve sync")) return -3;
    if (iterations)
        for (int k = 0; k < K; ++k) iterations[order_[size_t(k)]] = st_host_[size_t(k)].iters;
    return first_error_ == hipSuccess ? 0 : -3;
}

}  // namespace uaamg
