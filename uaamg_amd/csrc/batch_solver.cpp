// Batched UAAMG-PCG Poisson solve (see batch_solver.h).
#include "batch_solver.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <type_traits>

namespace uaamg {

namespace {
constexpr int kCoefs = 4;   // diag (with invdiag), x, y, z
}  // namespace

PoissonBatch::~PoissonBatch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    fb_dev_.clear();   // the fallback solvers run on stream_
    destroy_graphs();
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
    if (mailbox_) (void)hipHostFree(mailbox_);
}

void PoissonBatch::destroy_graphs() {
    for (Graph* g : {&g_init_, &g_first_, &g_iter_[0], &g_iter_[1]}) {
        if (g->e) (void)hipGraphExecDestroy(g->e);
        if (g->g) (void)hipGraphDestroy(g->g);
        g->e = nullptr;
        g->g = nullptr;
    }
    graphs_ready_ = false;
}

int PoissonBatch::check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (first_error_ == hipSuccess) {
            first_error_ = e;
            err_ = std::string(what) + ": " + hipGetErrorString(e);
        }
        return -3;
    }
    return 0;
}

int64_t PoissonBatch::ndof(int input_index) const {
    if (input_index < 0 || input_index >= nproblems()) return -1;
    return ndof_[size_t(rank_[size_t(input_index)])];
}

int PoissonBatch::init(std::vector<PoissonHierarchyHost>&& probs, const SolverOptions& opt) {
    opt_ = opt;
    const int K = int(probs.size());
    if (K == 0) {
        err_ = "empty batch";
        return -1;
    }
    if (opt_.smoother != kScheduledRelaxedJacobi && opt_.smoother != kWeightedJacobi) {
        err_ = "the batched solve runs the Jacobi-family smoothers";
        return -1;
    }
    if (opt_.precision != 0 && opt_.precision != 1) {
        err_ = "unsupported precision";
        return -1;
    }
    mixed_ = opt_.precision == 1;
    if (opt_.flags & 128) {   // FLIP's fallback needs each problem alone: keep the host hierarchies
        fb_host_ = probs;
        fb_dev_.resize(size_t(K));
    }
    use_graphs_ = !(opt_.flags & 2);
    block_max_leaves_ = (opt_.flags & 16) ? 0 : PoissonDevice::kTuneMaxLeaves;
    if (const char* env = std::getenv("UAAMG_BLOCK_LEAVES")) block_max_leaves_ = std::atoi(env);
    for (const auto& p : probs) {
        if (p.levels.empty() || p.levels[0].ndof == 0) {
            err_ = "a problem has no degrees of freedom";
            return -1;
        }
        if (p.levels.back().ndof > kMaxCoarsestDof) {
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
    auto hl = [&](int k, int l) -> const PoissonLevelHost& { return probs[size_t(order_[size_t(k)])].levels[size_t(l)]; };

    // leaf offsets per level; level 0 aligned to the reduction block
    std::vector<std::vector<int64_t>> off(static_cast<size_t>(L));
    std::vector<int64_t> nleaf(size_t(L), 0), nsmooth(size_t(L), 0);
    lv_.assign(size_t(L), Level{});
    for (int l = 0; l < L; ++l) {
        int64_t o = 0;
        for (int k = 0; k < K && nlev_[size_t(k)] > l; ++k) {
            off[size_t(l)].push_back(o);
            o += int64_t(hl(k, l).nleaf());
            if (l == 0) o = (o + kWavesPerBlock - 1) / kWavesPerBlock * kWavesPerBlock;
            if (nlev_[size_t(k)] > l + 1) nsmooth[size_t(l)] = o;
        }
        nleaf[size_t(l)] = o;
        int first = -1, count = 0;
        for (int k = 0; k < K; ++k)
            if (nlev_[size_t(k)] == l + 1) {
                if (first < 0) first = k;
                ++count;
            }
        lv_[size_t(l)].coarse_first = std::max(first, 0);
        lv_[size_t(l)].coarse_count = count;
    }
    if (nleaf[0] > int64_t(INT32_MAX) / 512) {
        err_ = "batch too large for 32-bit brick positions";
        return -1;
    }

    // host images of the merged tables
    struct HostLevel {
        std::vector<uint64_t> mask;
        std::vector<int32_t> nbr, nb27, slot, child, parent;
        std::vector<int64_t> dofbase;
        std::vector<float> pool[kCoefs + 1];   // diag, x, y, z, invdiag (shares the diag slots)
    };
    std::vector<HostLevel> H(static_cast<size_t>(L));
    for (int l = 0; l < L; ++l) {
        HostLevel& h = H[size_t(l)];
        const size_t n = size_t(nleaf[size_t(l)]);
        h.mask.assign(n * 8, 0);
        h.nbr.assign(n * 6, -1);
        h.nb27.assign(n * 27, -1);
        h.slot.assign(n * kSlotStride, -1);
        h.child.assign(n * 8, -1);
        h.parent.assign(n, -1);
        h.dofbase.assign(n, 0);
        for (int k = 0; k < int(off[size_t(l)].size()); ++k) {
            const PoissonLevelHost& s = hl(k, l);
            const int64_t o = off[size_t(l)][size_t(k)];
            const size_t m = s.nleaf();
            std::memcpy(&h.mask[size_t(o) * 8], s.mask.data(), m * 8 * sizeof(uint64_t));
            for (size_t j = 0; j < m; ++j) {
                h.dofbase[size_t(o) + j] = s.dofbase[j];
                for (int d = 0; d < 6; ++d) {
                    const int32_t nb = s.nbr[j * 6 + size_t(d)];
                    h.nbr[(size_t(o) + j) * 6 + size_t(d)] = nb >= 0 ? int32_t(nb + o) : -1;
                }
                for (int q = 0; q < 27; ++q) {
                    const int32_t nb = s.lookup.find({s.origin[j].x + 8 * (q / 9 - 1), s.origin[j].y + 8 * ((q / 3) % 3 - 1),
                                                      s.origin[j].z + 8 * (q % 3 - 1)});
                    h.nb27[(size_t(o) + j) * 27 + size_t(q)] = nb >= 0 ? int32_t(nb + o) : -1;
                }
                if (l > 0)
                    for (int b = 0; b < 8; ++b) {
                        const int32_t c = s.child[j * 8 + size_t(b)];
                        h.child[(size_t(o) + j) * 8 + size_t(b)] = c >= 0 ? int32_t(c + off[size_t(l - 1)][size_t(k)]) : -1;
                    }
            }
            // coefficient bricks: kept leaves, then this problem's default bricks
            const int kinds[kCoefs] = {kCoefDiag, kCoefX, kCoefY, kCoefZ};
            int32_t dflt_slot[kCoefs];
            for (int c = 0; c < kCoefs; ++c) {
                std::vector<float>& pool = h.pool[c];
                for (size_t j = 0; j < m; ++j)
                    if (s.keep[kinds[c]][j]) {
                        h.slot[(size_t(o) + j) * kSlotStride + size_t(kinds[c])] = int32_t(pool.size() / kLeafSize);
                        pool.insert(pool.end(), s.coef[kinds[c]].begin() + long(j * kLeafSize),
                                    s.coef[kinds[c]].begin() + long((j + 1) * kLeafSize));
                        if (c == 0)
                            h.pool[kCoefs].insert(h.pool[kCoefs].end(), s.invdiag.begin() + long(j * kLeafSize),
                                                  s.invdiag.begin() + long((j + 1) * kLeafSize));
                    }
                dflt_slot[c] = int32_t(pool.size() / kLeafSize);
                pool.insert(pool.end(), kLeafSize, s.default_coef(kinds[c]));
                if (c == 0) h.pool[kCoefs].insert(h.pool[kCoefs].end(), kLeafSize, s.default_invdiag());
            }
            for (size_t j = 0; j < m; ++j)
                for (int c = 0; c < kCoefs; ++c) {
                    int32_t& sl = h.slot[(size_t(o) + j) * kSlotStride + size_t(kinds[c])];
                    if (sl < 0) sl = dflt_slot[c];
                }
        }
    }
    for (int l = 0; l + 1 < L; ++l)
        for (size_t c = 0; c < size_t(nleaf[size_t(l + 1)]); ++c)
            for (int b = 0; b < 8; ++b) {
                const int32_t f = H[size_t(l + 1)].child[c * 8 + size_t(b)];
                if (f >= 0) H[size_t(l)].parent[size_t(f)] = int32_t(c * 8 + size_t(b));
            }
    for (int l = 0; l < L; ++l)
        complete_leaf_records(H[size_t(l)].slot.data(), H[size_t(l)].nbr.data(), H[size_t(l)].parent.data(),
                              size_t(nleaf[size_t(l)]));
    // coarsest systems, merged order
    std::vector<std::vector<int32_t>> cpos(static_cast<size_t>(K)), cnb(static_cast<size_t>(K));
    std::vector<std::vector<float>> ccoef(static_cast<size_t>(K));
    for (int k = 0; k < K; ++k) {
        const int c = nlev_[size_t(k)] - 1;
        coarsest_system(hl(k, c), off[size_t(c)][size_t(k)], cpos[size_t(k)], cnb[size_t(k)], ccoef[size_t(k)]);
        lv_[size_t(c)].coarse_max_ndof = std::max(lv_[size_t(c)].coarse_max_ndof, int(cpos[size_t(k)].size()));
    }

    // one arena for everything
    constexpr size_t kAl = 256;
    auto al = [](size_t b) { return (std::max<size_t>(b, 4) + kAl - 1) / kAl * kAl; };
    size_t bytes = 1 << 16;
    for (int l = 0; l < L; ++l) {
        const HostLevel& h = H[size_t(l)];
        const size_t n = size_t(nleaf[size_t(l)]);
        bytes += al(h.mask.size() * 8) + al(h.nbr.size() * 4) + (l >= 1 ? al(h.nb27.size() * kNb27Stride * 4) : 0) +
                 al(h.slot.size() * 4) +
                 al(h.child.size() * 4) + al(h.parent.size() * 4) + al(h.dofbase.size() * 8);
        for (const auto& p : h.pool) bytes += al(p.size() * 4);
        bytes += 4 * al(n * kLeafSize * 4);
        if (mixed_) bytes += 4 * al(n * kLeafSize * 2);
    }
    const size_t n0 = size_t(nleaf[0]);
    bytes += 7 * al(n0 * kLeafSize * 4);                      // x b r z q p p'
    bytes += al(n0 * 4) + al((size_t(K) + 1) * 4);            // leafprob, blk_begin
    bytes += al((n0 / kWavesPerBlock + 16) * 8);              // partials
    bytes += al(size_t(K) * sizeof(BatchPcg)) + al(64) + al(size_t(K) * sizeof(CoarsestView));
    for (int k = 0; k < K; ++k)
        bytes += al(cpos[size_t(k)].size() * 4) + al(cnb[size_t(k)].size() * 4) + al(ccoef[size_t(k)].size() * 4);

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
        const size_t n = size_t(nleaf[size_t(l)]);
        const PoissonLevelHost& s0 = hl(0, l);
        LevelView& v = d.v;
        v.nleaf = int32_t(n);
        v.level = l;
        v.term = s0.term;
        // read only where a neighbour leaf is absent, always times a zero value
        v.dflt_diag = s0.default_coef(kCoefDiag);
        v.dflt_off = s0.default_coef(kCoefX);
        v.dflt_invdiag = s0.default_invdiag();
        v.dflt_spai0 = s0.default_spai0_apply();
        v.mask = static_cast<const uint64_t*>(up(h.mask.data(), h.mask.size() * 8));
        v.nbr = static_cast<const int32_t*>(up(h.nbr.data(), h.nbr.size() * 4));
        std::vector<int32_t> nbrec;
        if (l >= 1) {
            nbrec.resize(h.nb27.size() * kNb27Stride);
            build_nb27_records(nbrec.data(), h.nb27.data(), h.slot.data(), h.parent.data(), n);
            v.nb27 = static_cast<const int32_t*>(up(nbrec.data(), nbrec.size() * 4));
            (void)check(hipStreamSynchronize(stream_), "upload sync");   // nbrec dies at scope end
        } else {
            v.nb27 = nullptr;
        }
        v.slot = static_cast<const int32_t*>(up(h.slot.data(), h.slot.size() * 4));
        v.dofbase = static_cast<const int64_t*>(up(h.dofbase.data(), h.dofbase.size() * 8));
        v.parent = static_cast<const int32_t*>(up(h.parent.data(), h.parent.size() * 4));
        v.child = static_cast<const int32_t*>(up(h.child.data(), h.child.size() * 4));
        v.diag = static_cast<const float*>(up(h.pool[0].data(), h.pool[0].size() * 4));
        v.cx = static_cast<const float*>(up(h.pool[1].data(), h.pool[1].size() * 4));
        v.cy = static_cast<const float*>(up(h.pool[2].data(), h.pool[2].size() * 4));
        v.cz = static_cast<const float*>(up(h.pool[3].data(), h.pool[3].size() * 4));
        v.invdiag = static_cast<const float*>(up(h.pool[kCoefs].data(), h.pool[kCoefs].size() * 4));
        v.spai0 = nullptr;
        d.sv = v;
        d.sv.nleaf = int32_t(nsmooth[size_t(l)]);
        d.X = static_cast<float*>(zeros(n * kLeafSize * 4));
        d.B = static_cast<float*>(zeros(n * kLeafSize * 4));
        d.T = static_cast<float*>(zeros(n * kLeafSize * 4));
        d.U = static_cast<float*>(zeros(n * kLeafSize * 4));
        if (mixed_) {
            d.Xh = static_cast<bf16*>(zeros(n * kLeafSize * 2));
            d.Bh = static_cast<bf16*>(zeros(n * kLeafSize * 2));
            d.Th = static_cast<bf16*>(zeros(n * kLeafSize * 2));
            d.Uh = static_cast<bf16*>(zeros(n * kLeafSize * 2));
        }
        // the host images die with this function: finish the copies first
        if (check(hipStreamSynchronize(stream_), "upload sync")) return -3;
    }
    x0_ = static_cast<float*>(zeros(n0 * kLeafSize * 4));
    b0_ = static_cast<float*>(zeros(n0 * kLeafSize * 4));
    r0_ = static_cast<float*>(zeros(n0 * kLeafSize * 4));
    z0_ = static_cast<float*>(zeros(n0 * kLeafSize * 4));
    q0_ = static_cast<float*>(zeros(n0 * kLeafSize * 4));
    p_[0] = static_cast<float*>(zeros(n0 * kLeafSize * 4));
    p_[1] = static_cast<float*>(zeros(n0 * kLeafSize * 4));
    {
        std::vector<int32_t> leafprob(n0, 0), blk(size_t(K) + 1, 0);
        off0_.assign(size_t(K), 0);
        for (int k = 0; k < K; ++k) {
            const int64_t o = off[0][size_t(k)];
            const int64_t e = k + 1 < K ? off[0][size_t(k + 1)] : int64_t(n0);
            off0_[size_t(k)] = o;
            for (int64_t j = o; j < e; ++j) leafprob[size_t(j)] = k;
            blk[size_t(k)] = int32_t(o / kWavesPerBlock);
        }
        blk[size_t(K)] = int32_t(n0 / kWavesPerBlock);
        leafprob_ = static_cast<const int32_t*>(up(leafprob.data(), leafprob.size() * 4));
        blk_begin_ = static_cast<const int32_t*>(up(blk.data(), blk.size() * 4));
        partials_ = static_cast<double*>(zeros((n0 / kWavesPerBlock + 16) * 8));
        st_ = static_cast<BatchPcg*>(zeros(size_t(K) * sizeof(BatchPcg)));
        seq_dev_ = static_cast<unsigned int*>(zeros(64));
        std::vector<CoarsestView> views(static_cast<size_t>(K));
        for (int k = 0; k < K; ++k) {
            CoarsestView& cv = views[size_t(k)];
            cv.ndof = int32_t(hl(k, nlev_[size_t(k)] - 1).ndof);
            cv.pos = static_cast<const int32_t*>(up(cpos[size_t(k)].data(), cpos[size_t(k)].size() * 4));
            cv.nb = static_cast<const int32_t*>(up(cnb[size_t(k)].data(), cnb[size_t(k)].size() * 4));
            cv.coef = static_cast<const float*>(up(ccoef[size_t(k)].data(), ccoef[size_t(k)].size() * 4));
            cv.max_iters = 10;
            cv.tol = FLT_EPSILON;
        }
        views_ = static_cast<CoarsestView*>(up(views.data(), views.size() * sizeof(CoarsestView)));
        if (check(hipStreamSynchronize(stream_), "upload sync")) return -3;
    }
    if (!p_[1] || !views_ || !st_ || !seq_dev_) {
        err_ = "arena too small";
        return -2;
    }
    // per-problem level-0 views (DOF scatter / gather) and the suffix of
    // problems whose only level is level 0
    sub0_.assign(size_t(K), LevelView{});
    const LevelView& v0 = lv_[0].v;
    int64_t c0_first = -1;
    for (int k = 0; k < K; ++k) {
        LevelView s = v0;
        const int64_t o = off0_[size_t(k)];
        s.nleaf = int32_t(hl(k, 0).nleaf());
        s.mask = v0.mask + o * 8;
        s.dofbase = v0.dofbase + o;
        sub0_[size_t(k)] = s;
        if (nlev_[size_t(k)] == 1 && c0_first < 0) c0_first = o;
    }
    if (c0_first >= 0) {
        c0_ = v0;
        c0_.nleaf = int32_t(int64_t(n0) - c0_first);
        c0_.mask = v0.mask + c0_first * 8;
        c0_block_ = c0_first / kWavesPerBlock;
    }
    if (check(hipHostMalloc(reinterpret_cast<void**>(&mailbox_), sizeof(HostMailbox),
                            hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc mailbox"))
        return -3;
    std::memset(const_cast<HostMailbox*>(mailbox_), 0, sizeof(HostMailbox));
    if (check(hipHostGetDevicePointer(reinterpret_cast<void**>(&mailbox_dev_), mailbox_, 0), "mailbox alias"))
        return -3;
    if (block_max_leaves_ > 0 && check(block_kernels_init(), "block kernel attributes")) return -3;
    st_host_.assign(size_t(K), BatchPcg{});
    // level-to-kernel mapping: the faster of block-visit and per-sweep launches
    // for each merged coarse level (PoissonDevice::tune_levels), unless
    // UAAMG_BLOCK_LEAVES fixes a leaf-count threshold
    use_block_.assign(size_t(L), 0);
    for (int l = 1; l < L; ++l)
        use_block_[size_t(l)] = lv_[size_t(l)].sv.nleaf > 0 && int64_t(lv_[size_t(l)].sv.nleaf) <= block_max_leaves_;
    if (block_max_leaves_ > 0 && !std::getenv("UAAMG_BLOCK_LEAVES")) tune_levels();
    return check(hipStreamSynchronize(stream_), "init sync");
}

void PoissonBatch::visit_pair(int l, bool block) {
    Level& d = lv_[size_t(l)];
    Level& c = lv_[size_t(l + 1)];
    const LevelView& v = d.sv;
    const bool weighted = opt_.smoother == kWeightedJacobi;
    auto wpre = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, i); };
    auto wpost = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, kSweeps - 1 - i); };
    for (int first = 1; first >= 0; --first) {
        if (block) {
            if (mixed_) {
                check(launch_block_pre(v, first, d.Xh, d.Bh, d.Th, c.Bh, wpre(0), wpre(1), wpre(2), stream_), "tune");
                check(launch_block_post(v, d.Th, c.Xh, d.Bh, d.Xh, wpost(0), wpost(1), wpost(2), stream_), "tune");
            } else {
                check(launch_block_pre(v, first, d.X, d.B, d.T, c.B, wpre(0), wpre(1), wpre(2), stream_), "tune");
                check(launch_block_post(v, d.T, c.X, d.B, d.X, wpost(0), wpost(1), wpost(2), stream_), "tune");
            }
        } else if (mixed_) {
            if (first) {
                check(launch_jacobi_first2(v, d.Bh, d.Th, wpre(0), wpre(1), stream_), "tune");
                check(launch_jacobi(v, d.Th, d.Bh, d.Uh, wpre(2), stream_), "tune");
            } else {
                check(launch_jacobi(v, d.Xh, d.Bh, d.Th, wpre(0), stream_), "tune");
                check(launch_jacobi(v, d.Th, d.Bh, d.Uh, wpre(1), stream_), "tune");
                check(launch_jacobi(v, d.Uh, d.Bh, d.Th, wpre(2), stream_), "tune");
            }
            check(launch_residual_restrict(v, d.Uh, d.Bh, c.Bh, stream_), "tune");
            check(launch_prolong_jacobi(v, d.Uh, c.Xh, d.Bh, d.Th, wpost(0), stream_), "tune");
            check(launch_jacobi(v, d.Th, d.Bh, d.Uh, wpost(1), stream_), "tune");
            check(launch_jacobi(v, d.Uh, d.Bh, d.Xh, wpost(2), stream_), "tune");
        } else {
            if (first) {
                check(launch_jacobi_first2(v, d.B, d.T, wpre(0), wpre(1), stream_), "tune");
                check(launch_apply(v, kModeJacobi, d.T, d.B, d.U, wpre(2), stream_), "tune");
            } else {
                check(launch_apply(v, kModeJacobi, d.X, d.B, d.T, wpre(0), stream_), "tune");
                check(launch_apply(v, kModeJacobi, d.T, d.B, d.U, wpre(1), stream_), "tune");
                check(launch_apply(v, kModeJacobi, d.U, d.B, d.T, wpre(2), stream_), "tune");
            }
            check(launch_residual_restrict(v, d.U, d.B, c.B, stream_), "tune");
            check(launch_prolong_jacobi(v, d.U, c.X, d.B, d.T, wpost(0), stream_), "tune");
            check(launch_apply(v, kModeJacobi, d.T, d.B, d.U, wpost(1), stream_), "tune");
            check(launch_apply(v, kModeJacobi, d.U, d.B, d.X, wpost(2), stream_), "tune");
        }
    }
}

// Times both mappings of every candidate level on the (all-zero) scratch
// vectors with HIP events and keeps the faster; zero in, zero out.
void PoissonBatch::tune_levels() {
    const int L = nlevels();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (check(hipEventCreate(&e0), "event") || check(hipEventCreate(&e1), "event")) return;
    constexpr int kReps = 8;
    for (int l = 1; l + 1 < L; ++l) {
        if (!use_block_[size_t(l)]) continue;
        float ms[2] = {0.f, 0.f};
        for (int variant = 0; variant < 2; ++variant) {
            visit_pair(l, variant == 0);
            check(hipEventRecord(e0, stream_), "event");
            for (int r = 0; r < kReps; ++r) visit_pair(l, variant == 0);
            check(hipEventRecord(e1, stream_), "event");
            check(hipEventSynchronize(e1), "event");
            check(hipEventElapsedTime(&ms[variant], e0, e1), "event");
        }
        use_block_[size_t(l)] = ms[0] < ms[1] ? 1 : 0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// ------------------------------------------------------------ W-cycle
// muCyclePreconditioner<2, true> (simd_vdb_poisson.cpp:1844-1977) on every
// problem of the batch at once: the problems whose coarsest level is l solve
// it here (first visit only: the second visit's identical zero-start solve is
// skipped, as in the single-problem cycle); the others smooth on the prefix
// view, recurse twice, and smooth again. dot_op >= 0 (level 0) forms <z, r>.
void PoissonBatch::precond(int l, float* X, const float* B, bool first, int dot_op) {
    Level& d = lv_[size_t(l)];
    if (first && d.coarse_count)
        check(launch_bcoarsest(views_ + d.coarse_first, d.coarse_count, d.coarse_max_ndof, B, X, stream_),
              "coarsest");
    const LevelView& v = d.sv;
    if (v.nleaf > 0) {
        const bool weighted = opt_.smoother == kWeightedJacobi;
        auto wpre = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, i); };
        auto wpost = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, kSweeps - 1 - i); };
        Level& c = lv_[size_t(l + 1)];
        if (l >= 1 && use_block_[size_t(l)]) {
            check(launch_block_pre(v, first, X, B, d.T, c.B, wpre(0), wpre(1), wpre(2), stream_), "block pre");
            precond(l + 1, c.X, c.B, true, -1);
            if (c.sv.nleaf > 0) precond(l + 1, c.X, c.B, false, -1);
            check(launch_block_post(v, d.T, c.X, B, X, wpost(0), wpost(1), wpost(2), stream_), "block post");
        } else {
            float* cur = X;
            auto next = [&]() { return cur == d.T ? d.U : d.T; };
            if (first) {
                float* dst = next();
                check(launch_jacobi_first2(v, B, dst, wpre(0), wpre(1), stream_), "first2");
                cur = dst;
                dst = next();
                check(launch_apply(v, kModeJacobi, cur, B, dst, wpre(2), stream_), "jacobi");
                cur = dst;
            } else {
                for (int i = 0; i < kSweeps; ++i) {
                    float* dst = next();
                    check(launch_apply(v, kModeJacobi, cur, B, dst, wpre(i), stream_), "jacobi");
                    cur = dst;
                }
            }
            check(launch_residual_restrict(v, cur, B, c.B, stream_), "residual restrict");
            precond(l + 1, c.X, c.B, true, -1);
            if (c.sv.nleaf > 0) precond(l + 1, c.X, c.B, false, -1);
            float* dst = next();
            check(launch_prolong_jacobi(v, cur, c.X, B, dst, wpost(0), stream_), "prolong jacobi");
            cur = dst;
            dst = next();
            check(launch_apply(v, kModeJacobi, cur, B, dst, wpost(1), stream_), "jacobi");
            cur = dst;
            if (l == 0 && dot_op >= 0)
                check(launch_jacobi_dot_partials(v, cur, B, X, wpost(2), partials_, stream_), "jacobi dot");
            else
                check(launch_apply(v, kModeJacobi, cur, B, X, wpost(2), stream_), "jacobi");
        }
    }
    if (l == 0 && dot_op >= 0) level0_dot(X, B, dot_op);
}

// The same cycle with the level vectors in BF16 (PoissonDevice::precond_mixed).
template <class TX, class TB>
void PoissonBatch::precond_mixed(int l, TX* X, const TB* B, bool first, int dot_op) {
    Level& d = lv_[size_t(l)];
    if (first && d.coarse_count)
        check(launch_bcoarsest(views_ + d.coarse_first, d.coarse_count, d.coarse_max_ndof, B, X, stream_),
              "coarsest");
    const LevelView& v = d.sv;
    if (v.nleaf > 0) {
        const bool weighted = opt_.smoother == kWeightedJacobi;
        auto wpre = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, i); };
        auto wpost = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, kSweeps - 1 - i); };
        Level& c = lv_[size_t(l + 1)];
        auto coarse = [&]() {
            precond_mixed<bf16, bf16>(l + 1, c.Xh, c.Bh, true, -1);
            if (c.sv.nleaf > 0) precond_mixed<bf16, bf16>(l + 1, c.Xh, c.Bh, false, -1);
        };
        bool done = false;
        if constexpr (std::is_same_v<TX, bf16>) {
            if (l >= 1 && use_block_[size_t(l)]) {
                check(launch_block_pre(v, first, X, B, d.Th, c.Bh, wpre(0), wpre(1), wpre(2), stream_), "block pre");
                coarse();
                check(launch_block_post(v, d.Th, c.Xh, B, X, wpost(0), wpost(1), wpost(2), stream_), "block post");
                done = true;
            }
        }
        if (!done) {
            bf16* cur = nullptr;
            auto next = [&]() { return cur == d.Th ? d.Uh : d.Th; };
            if (first) {
                bf16* dst = next();
                check(launch_jacobi_first2(v, B, dst, wpre(0), wpre(1), stream_), "first2");
                cur = dst;
                dst = next();
                check(launch_jacobi(v, cur, B, dst, wpre(2), stream_), "jacobi");
                cur = dst;
            } else {
                bf16* dst = next();
                check(launch_jacobi(v, X, B, dst, wpre(0), stream_), "jacobi");
                cur = dst;
                for (int i = 1; i < kSweeps; ++i) {
                    dst = next();
                    check(launch_jacobi(v, cur, B, dst, wpre(i), stream_), "jacobi");
                    cur = dst;
                }
            }
            check(launch_residual_restrict(v, cur, B, c.Bh, stream_), "residual restrict");
            coarse();
            bf16* dst = next();
            check(launch_prolong_jacobi(v, cur, c.Xh, B, dst, wpost(0), stream_), "prolong jacobi");
            cur = dst;
            dst = next();
            check(launch_jacobi(v, cur, B, dst, wpost(1), stream_), "jacobi");
            cur = dst;
            bool dotted = false;
            if constexpr (std::is_same_v<TX, float>) {
                if (l == 0 && dot_op >= 0) {
                    check(launch_jacobi_dot_partials(v, cur, B, X, wpost(2), partials_, stream_), "jacobi dot");
                    dotted = true;
                }
            }
            if (!dotted) check(launch_jacobi(v, cur, B, X, wpost(2), stream_), "jacobi");
        }
    }
    if constexpr (std::is_same_v<TX, float>) {
        if (l == 0 && dot_op >= 0) level0_dot(X, B, dot_op);
    }
}

// <z, r> partials of the problems whose only level is level 0 (their z came
// from the coarsest solve), then the per-problem finalize of every problem.
void PoissonBatch::level0_dot(float* X, const float* B, int dot_op) {
    if (c0_.nleaf > 0) {
        const size_t o = size_t(c0_block_) * kWavesPerBlock * kLeafSize;
        check(launch_dot_partials(c0_, X + o, B + o, partials_ + c0_block_, stream_), "dot");
    }
    check(launch_bfinalize(dot_op, nproblems(), blk_begin_, partials_, st_, stream_), "finalize");
}

// ------------------------------------------------------------ PCG pieces
//   init     r = b - A x, nu0, states
//   first    p0 = M r, rho, q = A p0, sigma, alpha, r -= alpha q, nu
//   iter[k]  z = M r, rho_new, beta, x += alpha p_k, p_k' = z + beta p_k,
//            q = A p_k', sigma, alpha, r -= alpha q, nu   (problems still active;
//            a problem that converged in the previous piece gets its last x update)
void PoissonBatch::issue_init() {
    const LevelView& v0 = lv_[0].v;
    const int K = nproblems();
    check(launch_apply(v0, kModeResidual, x0_, b0_, r0_, 0.f, stream_), "residual");
    check(launch_absmax_partials(v0, r0_, partials_, stream_), "absmax");
    check(launch_bfinalize(kBFinNu0, K, blk_begin_, partials_, st_, stream_), "finalize");
    check(launch_bstatus(K, st_, seq_dev_, mailbox_dev_, stream_), "status");
}

void PoissonBatch::issue_first() {
    const LevelView& v0 = lv_[0].v;
    const int K = nproblems();
    if (mixed_) precond_mixed<float, float>(0, p_[0], r0_, true, kBFinRho);
    else precond(0, p_[0], r0_, true, kBFinRho);
    check(launch_laplacian_dot_partials(v0, p_[0], q0_, partials_, stream_), "A p");
    check(launch_bfinalize(kBFinSigma, K, blk_begin_, partials_, st_, stream_), "finalize");
    check(launch_baxpy_absmax(v0, leafprob_, st_, q0_, x0_, p_[0], b0_, r0_, partials_, stream_), "axpy");
    check(launch_bfinalize(kBFinNuFirst, K, blk_begin_, partials_, st_, stream_), "finalize");
    check(launch_bstatus(K, st_, seq_dev_, mailbox_dev_, stream_), "status");
}

void PoissonBatch::issue_iter(int k) {
    const LevelView& v0 = lv_[0].v;
    const int K = nproblems();
    if (mixed_) precond_mixed<float, float>(0, z0_, r0_, true, kBFinRhoNew);
    else precond(0, z0_, r0_, true, kBFinRhoNew);
    check(launch_bupdate_laplacian_dot(v0, leafprob_, st_, z0_, p_[k], p_[k ^ 1], x0_, q0_, partials_, stream_),
          "update A p");
    check(launch_bfinalize(kBFinSigma, K, blk_begin_, partials_, st_, stream_), "finalize");
    check(launch_baxpy_absmax(v0, leafprob_, st_, q0_, x0_, p_[k ^ 1], b0_, r0_, partials_, stream_), "axpy");
    check(launch_bfinalize(kBFinNuIter, K, blk_begin_, partials_, st_, stream_), "finalize");
    check(launch_bstatus(K, st_, seq_dev_, mailbox_dev_, stream_), "status");
}

int PoissonBatch::capture(Graph& G, const std::function<void()>& body) {
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

int PoissonBatch::run(Graph& g, const std::function<void()>& body) {
    if (use_graphs_) return check(hipGraphLaunch(g.e, stream_), "graph launch");
    body();
    return first_error_ == hipSuccess ? 0 : -3;
}

int PoissonBatch::wait_active(int* active) {
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

int PoissonBatch::solve(float* const* x, const float* const* b, int* status, int* iterations) {
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
        if (!rc) rc = capture(g_iter_[0], [&] { issue_iter(0); });
        if (!rc) rc = capture(g_iter_[1], [&] { issue_iter(1); });
        if (rc) {
            destroy_graphs();
            return rc;
        }
        graphs_ready_ = true;
    }
    if (own_stream_ && check(caller_.follow(stream_), "order after the caller's stream")) return -3;
    const bool refine = mixed_ || (opt_.flags & 64);
    for (int k = 0; k < K; ++k) {
        BatchPcg& s = st_host_[size_t(k)];
        s = BatchPcg{};
        s.tol = opt_.relative_tolerance;
        s.max_iters = opt_.max_iterations;
        s.true_nu = -1.f;
        s.prev_true = INFINITY;
        s.replacements = refine ? 0 : -1;
        s.refine = refine ? 1 : 0;
    }
    check(hipMemcpyAsync(st_, st_host_.data(), size_t(K) * sizeof(BatchPcg), hipMemcpyHostToDevice, stream_),
          "upload state");
    for (int k = 0; k < K; ++k) {
        const int i = order_[size_t(k)];
        const size_t o = size_t(off0_[size_t(k)]) * kLeafSize;
        check(launch_scatter_dof(sub0_[size_t(k)], x[i], x0_ + o, stream_), "scatter x");
        check(launch_scatter_dof(sub0_[size_t(k)], b[i], b0_ + o, stream_), "scatter b");
    }
    int active = 0;
    if (run(g_init_, [&] { issue_init(); })) return -3;
    if (wait_active(&active)) return -3;
    // recurrences: the first from the initial residual, then (residual
    // replacement) restarts from the FP64 true residual of the problems whose
    // true residual misses the tolerance
    for (int round = 0; round <= kBMaxReplacements; ++round) {
        int k = 0;
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
                if (run(g_iter_[k], [&] { issue_iter(k); })) return -3;
                k ^= 1;
            }
            // problems that converged in the last piece: x += alpha p, then commit
            check(launch_bupdate_x(lv_[0].v, leafprob_, st_, p_[k], x0_, stream_), "update x");
            check(launch_bfinalize(-1, K, blk_begin_, partials_, st_, stream_), "commit");
        }
        if (!refine) break;
        check(launch_residual64_partials(lv_[0].v, x0_, b0_, r0_, partials_, stream_), "residual64");
        check(launch_bfinalize(kBFinReplace, K, blk_begin_, partials_, st_, stream_), "replace");
        check(launch_bstatus(K, st_, seq_dev_, mailbox_dev_, stream_), "status");
        if (wait_active(&active)) return -3;
        if (active == 0) break;
    }
    for (int kk = 0; kk < K; ++kk) {
        const int i = order_[size_t(kk)];
        const size_t o = size_t(off0_[size_t(kk)]) * kLeafSize;
        check(launch_gather_dof(sub0_[size_t(kk)], x0_ + o, x[i], stream_), "gather");
    }
    check(hipMemcpyAsync(st_host_.data(), st_, size_t(K) * sizeof(BatchPcg), hipMemcpyDeviceToHost, stream_),
          "download state");
    if (check(hipStreamSynchronize(stream_), "solve sync")) return -3;
    for (int kk = 0; kk < K; ++kk) {
        const int i = order_[size_t(kk)];
        const BatchPcg& s = st_host_[size_t(kk)];
        if (status) status[i] = s.state == kBSuccess ? 0 : 1;
        if (iterations) iterations[i] = s.iters;
    }
    last_.resize(size_t(K));
    for (int kk = 0; kk < K; ++kk) last_[size_t(order_[size_t(kk)])] = st_host_[size_t(kk)];
    mg_status_.assign(size_t(K), -1);
    mg_iters_.assign(size_t(K), -1);
    if (!fb_host_.empty())
        for (int i = 0; i < K; ++i)
            if (last_[size_t(i)].state != kBSuccess)
                if (int rc = fallback(i, x[i], b[i])) return rc;
    return first_error_ == hipSuccess ? 0 : -3;
}

// FLIP_vdb.cpp:3150-3155: when solveMultigridPCG fails, FLIP zeroes the
// pressure and calls solvePureMultigrid on the same hierarchy. The merged
// batch has no pure-multigrid cycle of its own, so the failed problem is
// solved alone by a single-problem solver on the batch's stream (made from the
// kept host hierarchy on the first failure, then reused).
int PoissonBatch::fallback(int i, float* x, const float* b) {
    std::unique_ptr<PoissonDevice>& d = fb_dev_[size_t(i)];
    if (!d) {
        SolverOptions o = opt_;
        o.stream = stream_;
        o.flags &= ~128;
        o.precision = 0;   // solvePureMultigrid runs the FP32 W-cycle
        d.reset(new PoissonDevice());
        PoissonHierarchyHost h = fb_host_[size_t(i)];
        if (int rc = d->init(std::move(h), o)) {
            err_ = "fallback solver of problem " + std::to_string(i) + ": " + d->error();
            d.reset();
            return rc;
        }
    }
    d->options().relative_tolerance = opt_.relative_tolerance;
    d->options().max_iterations = opt_.max_iterations;
    const int64_t n = ndof(i);
    if (check(hipMemsetAsync(x, 0, size_t(n) * sizeof(float), stream_), "zero the fallback's pressure")) return -3;
    int its = 0;
    const int rc = d->solve_mg_dev(x, b, &its);
    if (rc < 0) {
        err_ = "fallback of problem " + std::to_string(i) + ": " + d->error();
        return rc;
    }
    mg_status_[size_t(i)] = rc;
    mg_iters_[size_t(i)] = its;
    return 0;
}

}  // namespace uaamg
