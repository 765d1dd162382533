// Device-resident UAAMG-PCG Poisson solver (see poisson_solver.h).
#include "poisson_solver.h"

#include "build_timing.h"
#include "hierarchy_device.h"

#include <algorithm>
#include <tuple>
#include <mutex>
#include <map>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <utility>

namespace uaamg {

namespace {
constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
}  // namespace

// weights of the scheduled relaxed Jacobi smoother for n = 3 sweeps
// (muCyclePreconditioner, simd_vdb_poisson.cpp:1848-1866)
float scheduled_weight(int n, int i) {
    if (i >= n) return 6.0f / 7.0f;
    if (n == 1) return 6.0f / 7.0f;
    if (n == 2) return i == 0 ? 1.7319f : 0.5695f;
    static const float w3[3] = {2.2473f, 0.8571f, 0.5296f};
    return w3[i];
}
// muCycleIterative's schedule (:1982-1997): defaults 6/7, n >= 3 uses the 3-entry table
float scheduled_weight_iterative(int n, int i) {
    if (i >= n) return 6.0f / 7.0f;
    if (n == 2) return i == 0 ? 1.7319f : 0.5695f;
    if (n >= 3) {
        static const float w3[3] = {2.2473f, 0.8571f, 0.5296f};
        return i < 3 ? w3[i] : 6.0f / 7.0f;
    }
    return 6.0f / 7.0f;
}

DeviceArena::~DeviceArena() { release(); }

// The arena's hipMalloc / hipFree (device-wide synchronisation) stay apart
// from graph captures on other threads.
void DeviceArena::release() {
    if (base_) {
        std::lock_guard<std::recursive_mutex> lk(capture_lock());
        (void)hipFree(base_);
    }
    base_ = nullptr;
    cap_ = used_ = 0;
}

hipError_t DeviceArena::reserve(size_t bytes) {
    if (base_) return hipErrorInvalidValue;
    std::lock_guard<std::recursive_mutex> lk(capture_lock());
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&base_), bytes);
    if (e != hipSuccess) {
        base_ = nullptr;
        return e;
    }
    cap_ = bytes;
    used_ = 0;
    return hipSuccess;
}

void* DeviceArena::take(size_t bytes) {
    size_t b = align_up(bytes);
    if (used_ + b > cap_) return nullptr;
    void* p = base_ + used_;
    used_ += b;
    return p;
}

PoissonDevice::~PoissonDevice() { release_device(false); }

void PoissonDevice::release_device(bool keep_host) {
    if (keep_host && on_device()) (void)ensure_host_coefficients();
    // frees synchronise the device: kept apart from graph captures on other threads
    std::lock_guard<std::recursive_mutex> lk(capture_lock());
    if (stream_) (void)hipStreamSynchronize(stream_);
    destroy_graphs();
    if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
    own_stream_ = false;
    if (sc_host_) (void)hipHostFree(sc_host_);
    if (mailbox_) (void)hipHostFree(mailbox_);
    sc_host_ = nullptr;
    mailbox_ = nullptr;
    mailbox_dev_ = nullptr;
    nu_seq_ = 0;
    lv_.clear();
    x0_ = b0_ = r0_ = z0_ = q0_ = nullptr;
    p_[0] = p_[1] = nullptr;
    for (auto& s : staging_) s = nullptr;
    sc_ = nullptr;
    partials_ = nullptr;
    cv_ = CoarsestView{};
    coarsest_dof_of_ = nullptr;
    post_cg_ = false;
    for (void* p : spai0_bufs_) (void)hipFree(p);
    spai0_bufs_.clear();
    spai0_on_device_ = false;
    if (aux_) (void)hipFree(aux_);
    aux_ = nullptr;
    aux_bytes_ = 0;
    arena_.release();
}

void* PoissonDevice::aux(size_t bytes) {
    if (bytes <= aux_bytes_) return aux_;
    if (stream_) (void)hipStreamSynchronize(stream_);   // no queued work still reads the old buffer
    std::lock_guard<std::recursive_mutex> lk(capture_lock());
    if (aux_) (void)hipFree(aux_);
    aux_ = nullptr;
    aux_bytes_ = 0;
    const size_t grown = (bytes + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);   // whole MiB
    if (hipMalloc(&aux_, grown) != hipSuccess) {
        (void)hipGetLastError();
        aux_ = nullptr;
        return nullptr;
    }
    aux_bytes_ = grown;
    return aux_;
}

void PoissonDevice::destroy_graphs() {
    for (Graph* g : {&g_init_, &g_first_, &g_iter_[0], &g_iter_[1], &g_repl_[0], &g_repl_[1]}) {
        if (g->e) (void)hipGraphExecDestroy(g->e);
        if (g->g) (void)hipGraphDestroy(g->g);
        g->e = nullptr;
        g->g = nullptr;
    }
    graph_smoother_ = -1;
}

int PoissonDevice::capture(Graph& G, const std::function<void()>& body) {
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

// The W-cycle and the PCG iteration have a fixed launch sequence for a given
// smoother and SOR weights, so each piece is captured once and replayed:
//   init     r = b - A x, ||r||_inf
//   first    p0 = M r, rho, q = A p0, sigma, alpha, r -= alpha q, ||r||_inf
//   iter[k]  z = M r, rho_new, beta, x += alpha p_k, p_k' = z + beta p_k,
//            q = A p_k', sigma, alpha, r -= alpha q, ||r||_inf
// (iter[k] reads p from p_[k] and writes p_[k^1]).
void PoissonDevice::issue_init() {
    const LevelView& v0 = lv_[0].v;
    check(launch_apply(v0, kModeResidual, x0_, b0_, r0_, 0.f, stream_), "residual");
    check(launch_absmax(v0, r0_, sc_, partials_, mailbox_dev_, stream_), "absmax");
}

void PoissonDevice::issue_first() {
    const LevelView& v0 = lv_[0].v;
    precond(0, p_[0], r0_, true, kDotRho);
    check(launch_laplacian_dot(v0, p_[0], q0_, sc_, partials_, stream_), "A p");
    check(launch_axpy_absmax(v0, q0_, r0_, sc_, partials_, mailbox_dev_, stream_), "axpy");
}

void PoissonDevice::issue_iter(int k, bool replace) {
    const LevelView& v0 = lv_[0].v;
    precond(0, z0_, r0_, true, kDotRhoNew);
    check(launch_update_laplacian_dot(v0, z0_, p_[k], p_[k ^ 1], x0_, q0_, sc_, partials_, stream_), "update A p");
    if (replace)
        check(launch_replace64_absmax(v0, x0_, p_[k ^ 1], b0_, r0_, sc_, partials_, mailbox_dev_, stream_),
              "replace");
    else
        check(launch_axpy_absmax(v0, q0_, r0_, sc_, partials_, mailbox_dev_, stream_), "axpy");
}

int PoissonDevice::ensure_graphs() {
    if (graph_smoother_ == opt_.smoother && graph_sor_ == sor_w_ && graph_refine_ == refine_) return 0;
    destroy_graphs();
    int rc = capture(g_init_, [&] { issue_init(); });
    if (!rc) rc = capture(g_first_, [&] { issue_first(); });
    if (!rc) rc = capture(g_iter_[0], [&] { issue_iter(0); });
    if (!rc) rc = capture(g_iter_[1], [&] { issue_iter(1); });
    if (!rc && refine_) rc = capture(g_repl_[0], [&] { issue_iter(0, true); });
    if (!rc && refine_) rc = capture(g_repl_[1], [&] { issue_iter(1, true); });
    if (rc) {
        destroy_graphs();
        return rc;
    }
    graph_smoother_ = opt_.smoother;
    graph_sor_ = sor_w_;
    graph_refine_ = refine_;
    return 0;
}

// Replays a captured piece, or issues its launches directly when graphs are off.
int PoissonDevice::run(Graph& g, const std::function<void()>& body) {
    if (use_graphs_) return check(hipGraphLaunch(g.e, stream_), "graph launch");
    body();
    return first_error_ == hipSuccess ? 0 : -3;
}

// Waits for the next ||r||_inf reduction through the host-mapped mailbox; a
// stream error or an idle stream without the expected value ends the wait.
int PoissonDevice::wait_nu(float* nu) {
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

int PoissonDevice::check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (first_error_ == hipSuccess) {
            first_error_ = e;
            err_ = std::string(what) + ": " + hipGetErrorString(e);
        }
        return -3;
    }
    return 0;
}

int PoissonDevice::init(PoissonHierarchyHost&& h, const SolverOptions& opt) {
    host_ = std::move(h);
    opt_ = opt;
    if (host_.levels.empty()) {
        err_ = "empty hierarchy";
        return -1;
    }
    // SPAI0 coefficients only when that smoother is selected (ensure_spai0
    // builds and uploads them later when it is selected afterwards)
    if (opt_.smoother == kSpai0) {
        if (int rc = ensure_host_coefficients()) return rc;
        build_spai0(host_);
    }
    return acquire_device();
}

// Device memory of the hierarchy, the PCG vectors and the scratch: one arena.
size_t PoissonDevice::device_bytes() const {
    const int L = int(host_.levels.size());
    if (L == 0) return 0;
    size_t bytes = 0;
    auto add = [&](size_t b) { bytes += align_up(b); };
    const size_t n0 = host_.levels[0].nleaf();
    for (int l = 0; l < L; ++l) {
        const auto& hl = host_.levels[l];
        const size_t n = hl.nleaf();
        add(n * 8 * sizeof(uint64_t));
        add(n * 6 * sizeof(int32_t));
        add(n * kSlotStride * sizeof(int32_t));
        add(n * sizeof(int64_t));
        add(n * sizeof(int32_t));         // parent
        add(n * 8 * sizeof(int32_t));     // child
        if (l >= 1) add(n * 27 * kNb27Stride * sizeof(int32_t));   // neighbourhood records
        size_t nslot[kNumCoef] = {0, 0, 0, 0, 0};
        for (int k = 0; k < kNumCoef; ++k)
            for (size_t i = 0; i < n && !hl.keep[k].empty(); ++i) nslot[k] += hl.keep[k][i] ? 1 : 0;
        for (int k = 0; k < kNumCoef; ++k) add(std::max<size_t>(nslot[k], 1) * kLeafSize * sizeof(float));
        add(std::max<size_t>(nslot[kCoefDiag], 1) * kLeafSize * sizeof(float));   // invdiag
        add(4 * n * kLeafSize * sizeof(float));               // X, B, T, U
        if (opt_.precision == 1) add(4 * n * kLeafSize * sizeof(bf16));
    }
    add(10 * n0 * kLeafSize * sizeof(float));   // x0 b0 r0 p p' z0 q0 + 3 unit buffers
    add(3 * std::max<int64_t>(1, host_.levels[0].ndof) * sizeof(float));
    add(sizeof(PcgScalars));
    add((size_t(reduction_blocks(LevelView{int32_t(n0)})) + 16) * sizeof(double));
    const auto& hc = host_.levels.back();
    add(size_t(hc.ndof) * (sizeof(int32_t) * 7 + sizeof(float) * 8) + 1024);
    add(hc.nleaf() * kLeafSize * sizeof(int32_t));   // coarsest DOF at each brick position
    return bytes;
}

// Leaf records of a level (layout at kSlotDiag .. kSlotNbParent): the slots
// number each coefficient kind's kept bricks in leaf order; parent null
// leaves the parent fields -1.
static std::vector<int32_t> level_records(const PoissonLevelHost& h, const int32_t* parent) {
    const size_t n = h.nleaf();
    std::vector<int32_t> rec(n * kSlotStride, -1);
    int32_t count[kNumCoef] = {0, 0, 0, 0, 0};
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < kNumCoef; ++k)
            if (!h.keep[k].empty() && h.keep[k][i]) rec[i * kSlotStride + k] = count[k]++;
    complete_leaf_records(rec.data(), h.nbr.data(), parent, n);
    return rec;
}

int PoissonDevice::acquire_device() {
    if (on_device()) return 0;
    if (host_.levels.empty()) {
        err_ = "empty hierarchy";
        return -1;
    }
    first_error_ = hipSuccess;
    err_.clear();
    if (check(hipSetDevice(opt_.device), "hipSetDevice")) return -3;
    if (opt_.stream) {
        stream_ = opt_.stream;
    } else {
        std::lock_guard<std::recursive_mutex> lk(capture_lock());
        if (check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate")) return -3;
        own_stream_ = true;
    }
    const int L = int(host_.levels.size());
    sor_w_.assign(L, 1.2f);
    const size_t n0 = host_.levels[0].nleaf();
    bool tune_pending = false;
    StageClock clock;
    if (check(arena_.reserve(device_bytes()), "hipMalloc")) return -2;
    clock.lap("device: arena");

    lv_.resize(L);
    for (int l = 0; l < L; ++l)
        if (int rc = upload_level(host_.levels[l], lv_[l], true)) return rc;
    if (host_.dcoef) {
        // the packed arena arrays now hold the bricks; the build's full arrays go
        if (check(hipStreamSynchronize(stream_), "pack sync")) return -3;
        host_.dcoef.reset();
    }
    clock.lap("device: level upload");
    spai0_on_device_ = host_.spai0_ready;   // packed into the arena with the other coefficients
    // parent table of each finer level from the child table of the coarser one
    for (int l = 0; l + 1 < L; ++l) {
        const std::vector<int32_t> parent = level_parents(l);
        if (!parent.empty() && check(hipMemcpyAsync(const_cast<int32_t*>(lv_[l].v.parent), parent.data(),
                                                    parent.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_),
                                     "upload parent"))
            return -3;
        const std::vector<int32_t> rec = level_records(host_.levels[l], parent.data());
        if (!rec.empty() && check(hipMemcpyAsync(const_cast<int32_t*>(lv_[l].v.slot), rec.data(),
                                                 rec.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_),
                                  "upload leaf records"))
            return -3;
        std::vector<int32_t> nbrec;
        if (l >= 1 && lv_[l].v.nb27 && !parent.empty()) {
            const auto& hl = host_.levels[l];
            const size_t n = hl.nleaf();
            std::vector<int32_t> ids(n * 27);
            for (size_t i = 0; i < n; ++i)
                for (int k = 0; k < 27; ++k)
                    ids[i * 27 + size_t(k)] =
                        hl.lookup.find({hl.origin[i].x + 8 * (k / 9 - 1), hl.origin[i].y + 8 * ((k / 3) % 3 - 1),
                                        hl.origin[i].z + 8 * (k % 3 - 1)});
            nbrec.resize(n * 27 * kNb27Stride);
            build_nb27_records(nbrec.data(), ids.data(), rec.data(), parent.data(), n);
            if (check(hipMemcpyAsync(const_cast<int32_t*>(lv_[l].v.nb27), nbrec.data(), nbrec.size() * sizeof(int32_t),
                                     hipMemcpyHostToDevice, stream_),
                      "upload neighbourhood records"))
                return -3;
        }
        // the host vectors die at the end of this iteration
        if (check(hipStreamSynchronize(stream_), "upload parent")) return -3;
    }
    clock.lap("device: leaf records");
    auto vec = [&](size_t nleaf) { return static_cast<float*>(arena_.take(nleaf * kLeafSize * sizeof(float))); };
    x0_ = vec(n0);
    b0_ = vec(n0);
    r0_ = vec(n0);
    p_[0] = vec(n0);
    p_[1] = vec(n0);
    z0_ = vec(n0);
    q0_ = vec(n0);
    for (int i = 0; i < 3; ++i) staging_[i] = vec(n0);   // unit-op bricks
    sc_ = static_cast<PcgScalars*>(arena_.take(sizeof(PcgScalars)));
    partials_ = static_cast<double*>(arena_.take((size_t(reduction_blocks(lv_[0].v)) + 16) * sizeof(double)));
    if (!x0_ || !sc_ || !partials_) {
        err_ = "arena too small";
        return -2;
    }
    if (check(hipMemsetAsync(x0_, 0, 10 * n0 * kLeafSize * sizeof(float), stream_), "memset")) return -3;
    if (check(hipMemsetAsync(sc_, 0, sizeof(PcgScalars), stream_), "memset")) return -3;
    {
        std::lock_guard<std::recursive_mutex> lk(capture_lock());
        if (check(hipHostMalloc(reinterpret_cast<void**>(&sc_host_), sizeof(PcgScalars), hipHostMallocDefault),
                  "hipHostMalloc"))
            return -3;
        if (check(hipHostMalloc(reinterpret_cast<void**>(&mailbox_), sizeof(HostMailbox),
                                hipHostMallocMapped | hipHostMallocCoherent),
                  "hipHostMalloc mailbox"))
            return -3;
    }
    std::memset(const_cast<HostMailbox*>(mailbox_), 0, sizeof(HostMailbox));
    if (check(hipHostGetDevicePointer(reinterpret_cast<void**>(&mailbox_dev_), mailbox_, 0), "mailbox alias"))
        return -3;
    {
        const char* env = std::getenv("UAAMG_GRAPHS");
        use_graphs_ = !(env && env[0] == '0');
        env = std::getenv("UAAMG_FUSE");
        fuse_ = !(env && env[0] == '0');
        if (opt_.flags & 1) fuse_ = false;
        if (opt_.flags & 2) use_graphs_ = false;
        // level-to-kernel mapping: on request (flags bit 2 or
        // UAAMG_TAIL_LEAVES) the bottom pair of levels runs as one
        // single-workgroup launch when level L-2 is small. Off by default: at
        // 256^3 (level L-2 = 44 leaves) one workgroup takes 250 us per visit
        // pair against about 100 us for the per-sweep launches.
        int max_leaves = (opt_.flags & 4) ? 128 : 0;
        env = std::getenv("UAAMG_TAIL_LEAVES");
        if (env) max_leaves = std::atoi(env);
        tail_level_ = -1;
        if (max_leaves > 0 && L >= 3 && int64_t(host_.levels[L - 2].nleaf()) <= max_leaves) tail_level_ = L - 2;
        // level-to-kernel mapping of the coarse levels (1 .. L-2): each visit
        // as two block-visit launches or as the per-sweep kernels. By default
        // the faster of the two on this device, timed below for every level of
        // at most kTuneMaxLeaves leaves; UAAMG_BLOCK_LEAVES=n fixes it to
        // "levels of at most n leaves" instead; flags bit 4 switches the
        // block-visit kernels off.
        block_max_leaves_ = (opt_.flags & 16) ? 0 : kTuneMaxLeaves;
        env = std::getenv("UAAMG_BLOCK_LEAVES");
        const bool fixed = env != nullptr;
        if (env) block_max_leaves_ = std::atoi(env);
        use_block_.assign(size_t(L), 0);
        for (int l = 1; l + 1 < L; ++l)
            use_block_[size_t(l)] = int64_t(host_.levels[l].nleaf()) <= block_max_leaves_ ? 1 : 0;
        if (block_max_leaves_ > 0 && check(block_kernels_init(), "block kernel attributes")) return -3;
        tune_pending = block_max_leaves_ > 0 && !fixed;
    }
    if (opt_.precision != 0 && opt_.precision != 1) {
        err_ = "unsupported precision";
        return -1;
    }
    mixed_ = opt_.precision == 1;
    refine_ = mixed_ || (opt_.flags & 64);
    if (mixed_) {
        tail_level_ = -1;
        for (int l = 0; l < L; ++l) {
            const size_t n = host_.levels[l].nleaf() * kLeafSize;
            bf16** bufs[4] = {&lv_[l].Xh, &lv_[l].Bh, &lv_[l].Th, &lv_[l].Uh};
            for (bf16** b : bufs) {
                *b = static_cast<bf16*>(arena_.take(n * sizeof(bf16)));
                if (!*b && n) {
                    err_ = "arena too small";
                    return -2;
                }
                // rows without DOFs are never written by the sweeps
                if (n && check(hipMemsetAsync(*b, 0, n * sizeof(bf16), stream_), "memset")) return -3;
            }
        }
    }
    if (check(hipStreamSynchronize(stream_), "init sync")) return -3;
    clock.lap("device: vectors and options");
    if (int rc = build_coarsest()) return rc;
    clock.lap("device: coarsest system");
    if (tune_pending) tune_levels();
    if (check(hipStreamSynchronize(stream_), "init sync")) return -3;
    clock.lap("device: level mapping timing");
    return first_error_ == hipSuccess ? 0 : -3;
}

// One first visit and one repeated visit of level l's own work (pre-smoothing,
// residual, restriction, prolongation, post-smoothing; the coarser levels'
// work between them is left out), in the solver's precision, on the level's
// scratch vectors.
void PoissonDevice::visit_pair(int l, bool block) {
    const LevelView& v = lv_[l].v;
    DevLevel& d = lv_[l];
    DevLevel& c = lv_[l + 1];
    const bool weighted = opt_.smoother == kWeightedJacobi;
    auto wpre = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, i); };
    auto wpost = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, kSweeps - 1 - i); };
    // above the coarsest level, both mappings include the coarsest solve (the
    // block mapping runs it inside its post-visit launch)
    const bool cg = l + 2 == nlevels() && post_cg_;
    for (int first = 1; first >= 0; --first) {
        if (block) {
            if (mixed_) {
                check(launch_block_pre(v, first, d.Xh, d.Bh, d.Th, c.Bh, wpre(0), wpre(1), wpre(2), stream_), "tune");
                if (cg)
                    check(launch_block_post_cg(v, cv_, coarsest_dof_of_, c.Bh, c.Xh, d.Th, d.Bh, d.Xh, wpost(0),
                                               wpost(1), wpost(2), stream_),
                          "tune");
                else
                    check(launch_block_post(v, d.Th, c.Xh, d.Bh, d.Xh, wpost(0), wpost(1), wpost(2), stream_), "tune");
            } else {
                check(launch_block_pre(v, first, d.X, d.B, d.T, c.B, wpre(0), wpre(1), wpre(2), stream_), "tune");
                if (cg)
                    check(launch_block_post_cg(v, cv_, coarsest_dof_of_, c.B, c.X, d.T, d.B, d.X, wpost(0), wpost(1),
                                               wpost(2), stream_),
                          "tune");
                else
                    check(launch_block_post(v, d.T, c.X, d.B, d.X, wpost(0), wpost(1), wpost(2), stream_), "tune");
            }
            continue;
        }
        if (mixed_) {
            if (first) {
                check(launch_jacobi_first2(v, d.Bh, d.Th, wpre(0), wpre(1), stream_), "tune");
                check(launch_jacobi(v, d.Th, d.Bh, d.Uh, wpre(2), stream_), "tune");
            } else {
                check(launch_jacobi(v, d.Xh, d.Bh, d.Th, wpre(0), stream_), "tune");
                check(launch_jacobi(v, d.Th, d.Bh, d.Uh, wpre(1), stream_), "tune");
                check(launch_jacobi(v, d.Uh, d.Bh, d.Th, wpre(2), stream_), "tune");
            }
            check(launch_residual_restrict(v, d.Uh, d.Bh, c.Bh, stream_), "tune");
            if (cg) check(launch_coarsest_cg(cv_, c.Bh, c.Xh, stream_), "tune");
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
            if (cg) check(launch_coarsest_cg(cv_, c.B, c.X, stream_), "tune");
            check(launch_prolong_jacobi(v, d.U, c.X, d.B, d.T, wpost(0), stream_), "tune");
            check(launch_apply(v, kModeJacobi, d.T, d.B, d.U, wpost(1), stream_), "tune");
            check(launch_apply(v, kModeJacobi, d.U, d.B, d.X, wpost(2), stream_), "tune");
        }
    }
}

namespace {
// Timings of the two level mappings, kept per process and reused by handles
// whose level has about the same size (FLIP rebuilds its hierarchy every
// substep with nearly the same level sizes). Key: device, precision, level,
// and the leaf count in buckets of 2^(1/8) (about 9%). Either mapping gives
// bit-identical results, so a reused choice changes only the speed.
struct TuneKey {
    int device, mixed, level, bucket;
    bool operator<(const TuneKey& o) const {
        return std::tie(device, mixed, level, bucket) < std::tie(o.device, o.mixed, o.level, o.bucket);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<float, float>> g_tune_cache;
}  // namespace

// Times both kernel mappings of every candidate coarse level with HIP events
// on the solver's stream and keeps the faster. The scratch vectors hold zeros
// before and after (zero in, zero out), so no solve state is touched. Timings
// come from the process-wide cache when a level of about the same size was
// timed before (UAAMG_TUNE_CACHE=0 always times).
void PoissonDevice::tune_levels() {
    const int L = nlevels();
    tune_ms_[0].assign(size_t(L), -1.f);
    tune_ms_[1].assign(size_t(L), -1.f);
    const bool jacobi = opt_.smoother == kScheduledRelaxedJacobi || opt_.smoother == kWeightedJacobi;
    if (!fuse_ || !jacobi) return;
    const char* env = std::getenv("UAAMG_TUNE_CACHE");
    const bool cache = !(env && env[0] == '0');
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (check(hipEventCreate(&e0), "event") || check(hipEventCreate(&e1), "event")) return;
    constexpr int kReps = 8;
    for (int l = 1; l + 1 < L; ++l) {
        if (!use_block_[size_t(l)] || l + 1 == tail_level_) continue;
        const TuneKey key{opt_.device, mixed_ ? 1 : 0, l,
                          int(std::floor(std::log2(double(std::max<size_t>(host_.levels[size_t(l)].nleaf(), 1))) * 8))};
        if (cache) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune_cache.find(key);
            if (it != g_tune_cache.end()) {
                tune_ms_[0][size_t(l)] = it->second.first;
                tune_ms_[1][size_t(l)] = it->second.second;
                use_block_[size_t(l)] = it->second.first < it->second.second ? 1 : 0;
                continue;
            }
        }
        float ms[2] = {0.f, 0.f};
        for (int variant = 0; variant < 2; ++variant) {
            visit_pair(l, variant == 0);   // warm-up
            check(hipEventRecord(e0, stream_), "event");
            for (int r = 0; r < kReps; ++r) visit_pair(l, variant == 0);
            check(hipEventRecord(e1, stream_), "event");
            check(hipEventSynchronize(e1), "event");
            check(hipEventElapsedTime(&ms[variant], e0, e1), "event");
            ms[variant] /= kReps;
        }
        tune_ms_[0][size_t(l)] = ms[0];
        tune_ms_[1][size_t(l)] = ms[1];
        use_block_[size_t(l)] = ms[0] < ms[1] ? 1 : 0;
        if (cache) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune_cache[key] = {ms[0], ms[1]};
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

void PoissonDevice::init_host_only(PoissonHierarchyHost&& h, const SolverOptions& opt) {
    host_ = std::move(h);
    opt_ = opt;
}

// Builds the SPAI0 coefficients (init_SPAI0, simd_vdb_poisson.cpp:1113-1190)
// on first use and, on the device, uploads every level's kept SPAI0 bricks
// and the leaf records that point to them. Captured graphs are dropped (the
// next solve recaptures them with the new level views).
int PoissonDevice::ensure_spai0() {
    if (!host_.spai0_ready) {
        if (int rc = ensure_host_coefficients()) return rc;
        build_spai0(host_);
    }
    if (!on_device() || spai0_on_device_) return 0;
    for (int l = 0; l < nlevels(); ++l) {
        const PoissonLevelHost& h = host_.levels[size_t(l)];
        const size_t n = h.nleaf();
        size_t kept = 0;
        for (size_t i = 0; i < n; ++i) kept += h.keep[kCoefSpai0][i] ? 1 : 0;
        std::vector<float> buf(std::max<size_t>(kept, 1) * kLeafSize, 0.f);
        size_t j = 0;
        for (size_t i = 0; i < n; ++i)
            if (h.keep[kCoefSpai0][i]) {
                std::memcpy(&buf[j * kLeafSize], &h.coef[kCoefSpai0][i * kLeafSize], kLeafSize * sizeof(float));
                ++j;
            }
        void* d = nullptr;
        if (check(hipMalloc(&d, buf.size() * sizeof(float)), "spai0 alloc")) return -3;
        spai0_bufs_.push_back(d);
        if (check(hipMemcpyAsync(d, buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice, stream_),
                  "spai0 upload"))
            return -3;
        lv_[size_t(l)].v.spai0 = static_cast<const float*>(d);
        const std::vector<int32_t> parent = level_parents(l);
        const std::vector<int32_t> rec = level_records(h, parent.empty() ? nullptr : parent.data());
        if (!rec.empty() && check(hipMemcpyAsync(const_cast<int32_t*>(lv_[size_t(l)].v.slot), rec.data(),
                                                 rec.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_),
                                  "spai0 leaf records"))
            return -3;
        if (check(hipStreamSynchronize(stream_), "spai0 sync")) return -3;   // host buffers die here
    }
    spai0_on_device_ = true;
    destroy_graphs();
    return first_error_ == hipSuccess ? 0 : -3;
}

int PoissonDevice::ensure_host_coefficients() {
    bool any = false;
    for (const auto& l : host_.levels) any = any || l.coef_on_device;
    if (!any) return 0;
    if (host_.dcoef) {
        std::string e;
        if (!download_device_coefficients(host_, e)) {
            err_ = e;
            return -3;
        }
        return 0;
    }
    if (!on_device() || lv_.size() != host_.levels.size()) {
        err_ = "coefficient bricks are not available";
        return -1;
    }
    // from the packed arena arrays: kept bricks in slot order, the default elsewhere
    for (size_t l = 0; l < host_.levels.size(); ++l) {
        PoissonLevelHost& h = host_.levels[l];
        if (!h.coef_on_device) continue;
        const size_t n = h.nleaf();
        const LevelView& v = lv_[l].v;
        const float* src[5] = {v.diag, v.cx, v.cy, v.cz, v.invdiag};
        const int kinds[5] = {kCoefDiag, kCoefX, kCoefY, kCoefZ, kCoefDiag};
        for (int k = 0; k < 5; ++k) {
            const std::vector<uint8_t>& keep = h.keep[kinds[k]];
            size_t kept = 0;
            for (size_t i = 0; i < n; ++i) kept += keep[i] ? 1 : 0;
            BrickVec packed(std::max<size_t>(kept, 1) * kLeafSize);
            if (kept && check(hipMemcpyAsync(packed.data(), src[k], kept * kLeafSize * sizeof(float),
                                             hipMemcpyDeviceToHost, stream_),
                              "download bricks"))
                return -3;
            if (check(hipStreamSynchronize(stream_), "download bricks")) return -3;
            BrickVec& dst = k < 4 ? h.coef[kinds[k]] : h.invdiag;
            dst.resize(n * kLeafSize);
            const float dflt = k < 4 ? h.default_coef(kinds[k]) : h.default_invdiag();
            size_t j = 0;
            for (size_t i = 0; i < n; ++i) {
                float* out = &dst[i * kLeafSize];
                if (keep[i]) std::memcpy(out, &packed[(j++) * kLeafSize], kLeafSize * sizeof(float));
                else std::fill(out, out + kLeafSize, dflt);
            }
        }
        h.coef_on_device = false;
    }
    return 0;
}

// parent of every leaf of level l (coarse leaf * 8 + octant), empty on the coarsest level
std::vector<int32_t> PoissonDevice::level_parents(int l) const {
    if (l + 1 >= nlevels()) return {};
    const auto& coarse = host_.levels[size_t(l + 1)];
    std::vector<int32_t> parent(host_.levels[size_t(l)].nleaf(), -1);
    for (size_t c = 0; c < coarse.nleaf(); ++c)
        for (int b = 0; b < 8; ++b) {
            const int32_t f = coarse.child[c * 8 + size_t(b)];
            if (f >= 0) parent[size_t(f)] = int32_t(c * 8 + size_t(b));
        }
    return parent;
}

int PoissonDevice::upload_level(const PoissonLevelHost& h, DevLevel& d, bool spai0) {
    const size_t n = h.nleaf();
    LevelView& v = d.v;
    v.nleaf = int32_t(n);
    v.level = h.level;
    v.term = h.term;
    v.dflt_diag = h.default_coef(kCoefDiag);
    v.dflt_off = h.default_coef(kCoefX);
    v.dflt_invdiag = h.default_invdiag();
    v.dflt_spai0 = h.default_spai0_apply();
    auto up = [&](const void* src, size_t bytes) -> void* {
        void* p = arena_.take(std::max<size_t>(bytes, 4));
        if (p && bytes) (void)check(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, stream_), "upload");
        return p;
    };
    v.mask = static_cast<const uint64_t*>(up(h.mask.data(), n * 8 * sizeof(uint64_t)));
    v.nbr = static_cast<const int32_t*>(up(h.nbr.data(), n * 6 * sizeof(int32_t)));
    // neighbourhood records of the block-visit kernels (levels >= 1), filled
    // once the parents are known (acquire_device)
    v.nb27 = h.level >= 1 ? static_cast<const int32_t*>(arena_.take(std::max<size_t>(n * 27 * kNb27Stride * 4, 4)))
                          : nullptr;
    int32_t count[kNumCoef] = {0, 0, 0, 0, 0};
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < kNumCoef; ++k) count[k] += (!h.keep[k].empty() && h.keep[k][i]) ? 1 : 0;
    // parents are filled in once the coarser level is known (acquire_device)
    const std::vector<int32_t> slot = level_records(h, nullptr);
    v.slot = static_cast<const int32_t*>(up(slot.data(), slot.size() * sizeof(int32_t)));
    v.dofbase = static_cast<const int64_t*>(up(h.dofbase.data(), n * sizeof(int64_t)));
    v.parent = static_cast<const int32_t*>(arena_.take(std::max<size_t>(n * sizeof(int32_t), 4)));
    if (v.parent && check(hipMemsetAsync(const_cast<int32_t*>(v.parent), 0xff, std::max<size_t>(n * sizeof(int32_t), 4),
                                         stream_),
                          "memset"))
        return -3;
    if (!h.child.empty()) {
        v.child = static_cast<const int32_t*>(up(h.child.data(), n * 8 * sizeof(int32_t)));
    } else {
        v.child = static_cast<const int32_t*>(arena_.take(std::max<size_t>(n * 8 * sizeof(int32_t), 4)));
        if (v.child && check(hipMemsetAsync(const_cast<int32_t*>(v.child), 0xff,
                                            std::max<size_t>(n * 8 * sizeof(int32_t), 4), stream_),
                             "memset"))
            return -3;
    }
    // coefficient bricks of kept leaves only
    // at least one brick per array (zeros when no leaf keeps one), so a kernel
    // may read slot 0 unconditionally and select the default afterwards
    // bricks built on the device: packed there from the build's store
    const DeviceCoefStore::Level* dsrc =
        (h.coef_on_device && host_.dcoef && size_t(h.level) < host_.dcoef->levels.size() &&
         host_.dcoef->levels[size_t(h.level)].coef[0])
            ? &host_.dcoef->levels[size_t(h.level)]
            : nullptr;
    if (h.coef_on_device && !dsrc) {
        err_ = "level bricks are neither on the host nor in the device store";
        return -1;
    }
    int32_t* slot_d = nullptr;
    if (dsrc) {
        slot_d = static_cast<int32_t*>(aux(kNumCoef * std::max<size_t>(n, 1) * sizeof(int32_t)));
        if (!slot_d) {
            err_ = "device scratch allocation failed";
            return -2;
        }
        std::vector<int32_t> slots(kNumCoef * n, -1);
        for (int k = 0; k < 4; ++k) {
            int32_t j = 0;
            for (size_t i = 0; i < n; ++i)
                if (h.keep[k][i]) slots[size_t(k) * n + i] = j++;
        }
        if (n && check(hipMemcpyAsync(slot_d, slots.data(), 4 * n * sizeof(int32_t), hipMemcpyHostToDevice, stream_),
                       "upload slots"))
            return -3;
        (void)check(hipStreamSynchronize(stream_), "upload sync");   // the host table dies here
    }
    auto dpack = [&](const float* src, int kind) -> const float* {
        float* p = static_cast<float*>(arena_.take(std::max<size_t>(size_t(count[kind]), 1) * kLeafSize * sizeof(float)));
        if (!p) return nullptr;
        if (count[kind] == 0)
            (void)check(hipMemsetAsync(p, 0, kLeafSize * sizeof(float), stream_), "memset");
        else
            (void)check(launch_pack_bricks(int32_t(n), slot_d + size_t(kind) * n, src, p, stream_), "pack bricks");
        return p;
    };
    auto pack = [&](const BrickVec& src, int kind, const std::vector<uint8_t>& keep) -> const float* {
        if (size_t(count[kind]) == n && n > 0) {
            // every leaf keeps its brick: upload the array as it is
            const float* p = static_cast<const float*>(up(src.data(), n * kLeafSize * sizeof(float)));
            (void)check(hipStreamSynchronize(stream_), "upload sync");
            return p;
        }
        BrickVec buf(std::max<size_t>(size_t(count[kind]), 1) * kLeafSize);
        if (count[kind] == 0) std::fill(buf.begin(), buf.end(), 0.f);
        std::vector<int64_t> dst(n);
        int64_t j = 0;
        for (size_t i = 0; i < n; ++i) dst[i] = keep[i] ? j++ : -1;
        parallel_for(int64_t(n), [&](int64_t b, int64_t e) {
            for (int64_t i = b; i < e; ++i)
                if (dst[size_t(i)] >= 0)
                    std::memcpy(&buf[size_t(dst[size_t(i)]) * kLeafSize], &src[size_t(i) * kLeafSize],
                                kLeafSize * sizeof(float));
        });
        const float* p = static_cast<const float*>(up(buf.data(), buf.size() * sizeof(float)));
        // the host buffer must outlive the async copy
        (void)check(hipStreamSynchronize(stream_), "upload sync");
        return p;
    };
    if (dsrc) {
        v.diag = dpack(dsrc->coef[0], kCoefDiag);
        v.invdiag = dpack(dsrc->invdiag, kCoefDiag);
        v.cx = dpack(dsrc->coef[1], kCoefX);
        v.cy = dpack(dsrc->coef[2], kCoefY);
        v.cz = dpack(dsrc->coef[3], kCoefZ);
        if (!v.diag || !v.invdiag || !v.cx || !v.cy || !v.cz) {
            err_ = "arena too small";
            return -2;
        }
    } else {
        v.diag = pack(h.coef[kCoefDiag], kCoefDiag, h.keep[kCoefDiag]);
        v.invdiag = pack(h.invdiag, kCoefDiag, h.keep[kCoefDiag]);
        v.cx = pack(h.coef[kCoefX], kCoefX, h.keep[kCoefX]);
        v.cy = pack(h.coef[kCoefY], kCoefY, h.keep[kCoefY]);
        v.cz = pack(h.coef[kCoefZ], kCoefZ, h.keep[kCoefZ]);
    }
    if (spai0 && host_.spai0_ready) v.spai0 = pack(h.coef[kCoefSpai0], kCoefSpai0, h.keep[kCoefSpai0]);
    d.X = static_cast<float*>(arena_.take(n * kLeafSize * sizeof(float)));
    d.B = static_cast<float*>(arena_.take(n * kLeafSize * sizeof(float)));
    d.T = static_cast<float*>(arena_.take(n * kLeafSize * sizeof(float)));
    d.U = static_cast<float*>(arena_.take(n * kLeafSize * sizeof(float)));
    if (!d.X || !d.B || !d.T || !d.U) {
        err_ = "arena too small";
        return -2;
    }
    if (n) {
        if (check(hipMemsetAsync(d.X, 0, n * kLeafSize * sizeof(float), stream_), "memset")) return -3;
        if (check(hipMemsetAsync(d.B, 0, n * kLeafSize * sizeof(float), stream_), "memset")) return -3;
        if (check(hipMemsetAsync(d.T, 0, n * kLeafSize * sizeof(float), stream_), "memset")) return -3;
        if (check(hipMemsetAsync(d.U, 0, n * kLeafSize * sizeof(float), stream_), "memset")) return -3;
    }
    if (check(hipStreamSynchronize(stream_), "upload sync")) return -3;
    return first_error_ == hipSuccess ? 0 : -3;
}

void coarsest_system(const PoissonLevelHost& l, int64_t leaf_offset, std::vector<int32_t>& pos,
                     std::vector<int32_t>& nb, std::vector<float>& coef) {
    const int64_t nd = l.ndof;
    pos.assign(size_t(nd), 0);
    nb.assign(size_t(nd) * 6, -1);
    coef.assign(size_t(nd) * 8, 0.f);
    // DOF index lookup on the coarsest level
    auto dof_at = [&](const Coord& g) -> int64_t {
        int32_t li = l.lookup.find(leaf_origin_of(g));
        if (li < 0) return -1;
        const int off = voxel_offset(g.x & 7, g.y & 7, g.z & 7);
        const uint64_t* m = &l.mask[size_t(li) * 8];
        if (!mask_test(m, off)) return -1;
        int64_t k = l.dofbase[li];
        for (int o = 0; o < off; ++o) k += mask_test(m, o);
        return k;
    };
    auto coef_at = [&](int kind, const Coord& g) -> float {
        int32_t li = l.lookup.find(leaf_origin_of(g));
        if (li < 0) return l.default_coef(kind);
        return l.coef[kind][size_t(li) * kLeafSize + voxel_offset(g.x & 7, g.y & 7, g.z & 7)];
    };
    for (size_t i = 0; i < l.nleaf(); ++i) {
        int64_t k = l.dofbase[i];
        const Coord o = l.origin[i];
        for (int off = 0; off < kLeafSize; ++off) {
            if (!mask_test(&l.mask[i * 8], off)) continue;
            const Coord g{o.x + (off >> 6), o.y + ((off >> 3) & 7), o.z + (off & 7)};
            pos[k] = int32_t((int64_t(i) + leaf_offset) * kLeafSize + off);
            const float dg = l.coef[kCoefDiag][i * kLeafSize + off];
            coef[k * 8 + 0] = dg;
            for (int c = 0; c < 3; ++c) {
                Coord up = g, dn = g;
                up[c] += 1;
                dn[c] -= 1;
                nb[k * 6 + 2 * c] = int32_t(dof_at(up));
                nb[k * 6 + 2 * c + 1] = int32_t(dof_at(dn));
                coef[k * 8 + 1 + 2 * c] = coef_at(kCoefX + c, up);
                coef[k * 8 + 2 + 2 * c] = coef_at(kCoefX + c, g);
            }
            coef[k * 8 + 7] = dg != 0.f ? 1.0f / dg : 1.0f;
            ++k;
        }
    }
}

int PoissonDevice::build_coarsest() {
    const PoissonLevelHost& l = host_.levels.back();
    const int64_t nd = l.ndof;
    if (nd > kMaxCoarsestDof) {
        err_ = "coarsest level exceeds the single-workgroup solver size";
        return -1;
    }
    std::vector<int32_t> pos, nb;
    std::vector<float> coef;
    coarsest_system(l, 0, pos, nb, coef);
    auto up = [&](const void* src, size_t bytes) -> void* {
        void* p = arena_.take(std::max<size_t>(bytes, 4));
        if (p && bytes) (void)check(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, stream_), "upload");
        return p;
    };
    cv_.ndof = int32_t(nd);
    cv_.pos = static_cast<const int32_t*>(up(pos.data(), pos.size() * sizeof(int32_t)));
    cv_.nb = static_cast<const int32_t*>(up(nb.data(), nb.size() * sizeof(int32_t)));
    cv_.coef = static_cast<const float*>(up(coef.data(), coef.size() * sizeof(float)));
    cv_.max_iters = 10;
    cv_.tol = FLT_EPSILON;
    // the inverse of pos, for the post-visit kernel that takes the coarsest
    // solve in (UAAMG_POST_CG=0 keeps them apart; the results are the same)
    std::vector<int32_t> dof_of(size_t(l.nleaf()) * kLeafSize, -1);
    for (int64_t d = 0; d < nd; ++d) dof_of[size_t(pos[size_t(d)])] = int32_t(d);
    coarsest_dof_of_ = static_cast<const int32_t*>(up(dof_of.data(), dof_of.size() * sizeof(int32_t)));
    const char* env = std::getenv("UAAMG_POST_CG");
    post_cg_ = nd > 0 && coarsest_dof_of_ && coarsest_cg_threads() == 512 && !(env && env[0] == '0');
    if (check(hipStreamSynchronize(stream_), "coarsest upload")) return -3;
    return 0;
}

void PoissonDevice::smooth(int l, float*& cur, float*& oth, const float* B, int i, bool pre, bool iterative) {
    const LevelView& v = lv_[l].v;
    const int n = kSweeps;
    switch (opt_.smoother) {
        default:
        case kScheduledRelaxedJacobi: {
            const int idx = pre ? i : n - 1 - i;
            const float w = iterative ? scheduled_weight_iterative(n, idx) : scheduled_weight(n, idx);
            check(launch_apply(v, kModeJacobi, cur, B, oth, w, stream_), "jacobi");
            std::swap(cur, oth);
            break;
        }
        case kWeightedJacobi:
            check(launch_apply(v, kModeJacobi, cur, B, oth, 6.0f / 7.0f, stream_), "jacobi");
            std::swap(cur, oth);
            break;
        case kRedBlackGaussSeidel: {
            if (iterative) sor_w_[l] = 1.0f;
            const float w = sor_w_[l];
            const int first = pre ? kModeRedGS : kModeBlackGS, second = pre ? kModeBlackGS : kModeRedGS;
            check(launch_apply(v, first, cur, B, cur, w, stream_), "rbgs");
            check(launch_apply(v, second, cur, B, cur, w, stream_), "rbgs");
            break;
        }
        case kSpai0:
            check(launch_apply(v, kModeSpai0, cur, B, oth, 0.f, stream_), "spai0");
            std::swap(cur, oth);
            break;
    }
}

// muCyclePreconditioner<2, skip_first> (simd_vdb_poisson.cpp:1844-1977).
// dot_target >= 0 (level 0 only) also forms <X, B> into rho / rho_new.
void PoissonDevice::precond(int l, float* X, const float* B, bool skip_first, int dot_target) {
    const int L = nlevels();
    const LevelView& v = lv_[l].v;
    if (l == L - 1) {
        check(launch_coarsest_cg(cv_, B, X, stream_), "coarsest");
        if (dot_target >= 0) check(launch_dot(v, X, B, sc_, partials_, dot_target, stream_), "dot");
        return;
    }
    const bool jacobi = opt_.smoother == kScheduledRelaxedJacobi || opt_.smoother == kWeightedJacobi;
    if (mixed_ && fuse_ && jacobi) {
        precond_mixed<float, float>(l, X, B, skip_first, dot_target);
        return;
    }
    if (fuse_ && jacobi) {
        precond_fused(l, X, B, skip_first, dot_target);
        return;
    }
    float* cur = X;
    float* oth = lv_[l].T;
    int start = 0;
    if (skip_first) {
        switch (opt_.smoother) {
            default:
            case kScheduledRelaxedJacobi:
                check(launch_first_sweep(v, 0, B, oth, scheduled_weight(kSweeps, 0), stream_), "first");
                std::swap(cur, oth);
                break;
            case kWeightedJacobi:
                check(launch_first_sweep(v, 0, B, oth, 6.0f / 7.0f, stream_), "first");
                std::swap(cur, oth);
                break;
            case kRedBlackGaussSeidel:
                check(launch_first_sweep(v, 2, B, cur, sor_w_[l], stream_), "first");
                check(launch_apply(v, kModeBlackGS, cur, B, cur, sor_w_[l], stream_), "first black");
                break;
            case kSpai0:
                check(launch_first_sweep(v, 1, B, oth, 0.f, stream_), "first");
                std::swap(cur, oth);
                break;
        }
        start = 1;
    }
    for (int i = start; i < kSweeps; ++i) smooth(l, cur, oth, B, i, true, false);
    DevLevel& c = lv_[l + 1];
    if (fuse_) {
        check(launch_residual_restrict(v, cur, B, c.B, stream_), "residual restrict");
    } else {
        check(launch_apply(v, kModeResidual, cur, B, oth, 0.f, stream_), "residual");
        check(launch_restrict(c.v, oth, c.B, stream_), "restrict");
    }
    precond(l + 1, c.X, c.B, true, -1);
    // mu = 2: the second visit of the coarsest level repeats an identical
    // zero-guess solve, so it is skipped; other levels cycle again
    if (l + 1 != L - 1) precond(l + 1, c.X, c.B, false, -1);
    check(launch_prolong_add(v, c.X, cur, 1.0f, stream_), "prolong");
    for (int i = 0; i < kSweeps; ++i) smooth(l, cur, oth, B, i, false, false);
    if (dot_target >= 0) check(launch_dot(v, X, B, sc_, partials_, dot_target, stream_), "dot");
}

// The Jacobi-family W-cycle visit with fused passes: the first two sweeps from
// zero in one kernel, residual + restriction in one kernel, prolongation inside
// the first post-smoothing sweep, and the level-0 dot inside the last sweep.
// Sweeps rotate through T and U and the last one writes X, so the result lands
// in the caller's vector as with the reference's even number of pointer swaps.
void PoissonDevice::precond_fused(int l, float* X, const float* B, bool skip_first, int dot_target) {
    const int L = nlevels();
    const LevelView& v = lv_[l].v;
    DevLevel& d = lv_[l];
    const bool weighted = opt_.smoother == kWeightedJacobi;
    auto wpre = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, i); };
    auto wpost = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, kSweeps - 1 - i); };
    if (l >= 1 && l + 1 != tail_level_ && use_block_[size_t(l)]) {
        // two launches per visit: pre-smoothing + residual + restriction, then
        // prolongation + post-smoothing (the coarser visits run in between)
        DevLevel& c = lv_[l + 1];
        check(launch_block_pre(v, skip_first, X, B, d.T, c.B, wpre(0), wpre(1), wpre(2), stream_), "block pre");
        if (l + 1 == L - 1 && post_cg_) {
            check(launch_block_post_cg(v, cv_, coarsest_dof_of_, c.B, c.X, d.T, B, X, wpost(0), wpost(1), wpost(2),
                                       stream_),
                  "coarsest + block post");
            return;
        }
        precond(l + 1, c.X, c.B, true, -1);
        if (l + 1 != L - 1) precond(l + 1, c.X, c.B, false, -1);
        check(launch_block_post(v, d.T, c.X, B, X, wpost(0), wpost(1), wpost(2), stream_), "block post");
        return;
    }
    float* cur = X;
    auto next = [&]() { return cur == d.T ? d.U : d.T; };
    if (skip_first) {
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
    DevLevel& c = lv_[l + 1];
    check(launch_residual_restrict(v, cur, B, c.B, stream_), "residual restrict");
    if (l + 1 == tail_level_) {
        // both mu = 2 visits of level L-2 and their coarsest solves, one launch
        TailArgs ta{};
        ta.L = c.v;
        ta.B = c.B;
        ta.X = c.X;
        ta.T = c.T;
        ta.U = c.U;
        ta.Bc = lv_[L - 1].B;
        ta.Xc = lv_[L - 1].X;
        ta.C = cv_;
        for (int i = 0; i < kSweeps; ++i) {
            ta.wpre[i] = wpre(i);
            ta.wpost[i] = wpost(i);
        }
        ta.nvisits = 2;
        ta.first_skip = 1;
        check(launch_tail(ta, stream_), "tail");
    } else {
        precond(l + 1, c.X, c.B, true, -1);
        if (l + 1 != L - 1) precond(l + 1, c.X, c.B, false, -1);
    }
    {
        float* dst = next();
        check(launch_prolong_jacobi(v, cur, c.X, B, dst, wpost(0), stream_), "prolong jacobi");
        cur = dst;
        dst = next();
        check(launch_apply(v, kModeJacobi, cur, B, dst, wpost(1), stream_), "jacobi");
        cur = dst;
    }
    if (dot_target >= 0)
        check(launch_jacobi_dot(v, cur, B, X, wpost(2), sc_, partials_, dot_target, stream_), "jacobi dot");
    else
        check(launch_apply(v, kModeJacobi, cur, B, X, wpost(2), stream_), "jacobi");
}

// The fused Jacobi-family W-cycle with its level vectors stored in BF16: the
// same passes as precond_fused, every value computed in FP32 and rounded to
// BF16 where it is stored. Level 0 reads the FP32 residual B and its last
// sweep writes the FP32 X; below level 0 (TX = TB = bf16) every vector of the
// cycle is BF16.
template <class TX, class TB>
void PoissonDevice::precond_mixed(int l, TX* X, const TB* B, bool skip_first, int dot_target) {
    const int L = nlevels();
    const LevelView& v = lv_[l].v;
    DevLevel& d = lv_[l];
    if (l == L - 1) {
        check(launch_coarsest_cg(cv_, B, X, stream_), "coarsest");
        if constexpr (std::is_same_v<TX, float>) {
            if (dot_target >= 0) check(launch_dot(v, X, B, sc_, partials_, dot_target, stream_), "dot");
        }
        return;
    }
    const bool weighted = opt_.smoother == kWeightedJacobi;
    auto wpre = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, i); };
    auto wpost = [&](int i) { return weighted ? 6.0f / 7.0f : scheduled_weight(kSweeps, kSweeps - 1 - i); };
    DevLevel& c = lv_[l + 1];
    auto coarse = [&]() {
        precond_mixed<bf16, bf16>(l + 1, c.Xh, c.Bh, true, -1);
        if (l + 1 != L - 1) precond_mixed<bf16, bf16>(l + 1, c.Xh, c.Bh, false, -1);
    };
    if constexpr (std::is_same_v<TX, bf16>) {
        if (l >= 1 && use_block_[size_t(l)]) {
            check(launch_block_pre(v, skip_first, X, B, d.Th, c.Bh, wpre(0), wpre(1), wpre(2), stream_), "block pre");
            if (l + 1 == L - 1 && post_cg_) {
                check(launch_block_post_cg(v, cv_, coarsest_dof_of_, c.Bh, c.Xh, d.Th, B, X, wpost(0), wpost(1),
                                           wpost(2), stream_),
                      "coarsest + block post");
                return;
            }
            coarse();
            check(launch_block_post(v, d.Th, c.Xh, B, X, wpost(0), wpost(1), wpost(2), stream_), "block post");
            return;
        }
    }
    bf16* cur = nullptr;
    auto next = [&]() { return cur == d.Th ? d.Uh : d.Th; };
    if (skip_first) {
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
    {
        bf16* dst = next();
        check(launch_prolong_jacobi(v, cur, c.Xh, B, dst, wpost(0), stream_), "prolong jacobi");
        cur = dst;
        dst = next();
        check(launch_jacobi(v, cur, B, dst, wpost(1), stream_), "jacobi");
        cur = dst;
    }
    if constexpr (std::is_same_v<TX, float>) {
        if (dot_target >= 0) {
            check(launch_jacobi_dot(v, cur, B, X, wpost(2), sc_, partials_, dot_target, stream_), "jacobi dot");
            return;
        }
    }
    check(launch_jacobi(v, cur, B, X, wpost(2), stream_), "jacobi");
}

// muCycleIterative<2> (simd_vdb_poisson.cpp:1979-2110)
void PoissonDevice::mg_cycle(int l, float* X, const float* B) {
    const int L = nlevels();
    float* cur = X;
    float* oth = lv_[l].T;
    if (l == L - 1) {
        for (int i = 0; i < 10 * kSweeps; ++i) {
            if (opt_.smoother == kScheduledRelaxedJacobi) {
                check(launch_apply(lv_[l].v, kModeJacobi, cur, B, oth, scheduled_weight_iterative(kSweeps, i),
                                   stream_),
                      "jacobi");
                std::swap(cur, oth);
            } else {
                smooth(l, cur, oth, B, 0, true, true);
            }
        }
        return;
    }
    const LevelView& v = lv_[l].v;
    for (int i = 0; i < kSweeps; ++i) smooth(l, cur, oth, B, i, true, true);
    check(launch_apply(v, kModeResidual, cur, B, oth, 0.f, stream_), "residual");
    DevLevel& c = lv_[l + 1];
    check(launch_restrict(c.v, oth, c.B, stream_), "restrict");
    for (int mu = 0; mu < 2; ++mu) mg_cycle(l + 1, c.X, c.B);
    check(launch_prolong_add(v, c.X, cur, 0.5f, stream_), "prolong");
    for (int i = 0; i < kSweeps; ++i) smooth(l, cur, oth, B, i, false, true);
}

int PoissonDevice::follow_caller() {
    if (!own_stream_) return 0;
    return check(caller_.follow(stream_), "order after the caller's stream");
}

// solveMultigridPCG (simd_vdb_poisson.cpp:2155-2226). One captured piece per
// iteration runs the reference's lines 13-16 of iteration k and lines 6-8 of
// iteration k+1; the host reads ||r||_inf from the mailbox and applies the
// convergence and failure tests of lines 9-12 between pieces.
int PoissonDevice::solve_pcg_dev(float* x_dof, const float* b_dof, int* iterations) {
    if (!on_device()) return -4;
    if (first_error_ != hipSuccess) return -3;
    if (opt_.smoother == kSpai0 && ensure_spai0()) return -3;
    if (use_graphs_ && ensure_graphs()) return -3;
    if (follow_caller()) return -3;
    const LevelView& v0 = lv_[0].v;
    int it = 0;
    if (iterations) *iterations = 0;
    check(launch_scatter_dof(v0, x_dof, x0_, stream_), "scatter x");
    check(launch_scatter_dof(v0, b_dof, b0_, stream_), "scatter b");
    float nu;
    if (run(g_init_, [&] { issue_init(); })) return -3;
    if (wait_nu(&nu)) return -3;
    const float numax = opt_.relative_tolerance * nu;
    int status = kFailed;
    true_nu_ = -1.f;
    replacements_ = refine_ ? 0 : -1;
    bool nu_true = false;
    float prev = INFINITY;   // true residual at the last restart
    int restarts = 0;
    // a recurrence whose true residual stops halving ends for a restart when
    // the restart rule below would take one
    auto restart_below = [&] { return refine_ && restarts < kMaxReplacements ? 0.5f * prev : 0.f; };
    if (nu <= numax) {
        status = kSuccess;
    } else if (opt_.max_iterations > 0) {
        status = pcg_iterate(nu, numax, it, nu_true, restart_below(), refine_);
        if (status < 0) return status;
    }
    if (refine_ && status == kSuccess) {
        // FP64 residual replacement. Near convergence the first
        // recurrence's steps already replace r by the true residual
        // (pcg_iterate), ending it for a restart when that stops halving;
        // when the recursive residual met the tolerance without that,
        // recompute r = b - A x in FP64. While the true residual misses the
        // tolerance, restart the PCG recurrence from x on it (the iteration
        // count continues; restarted recurrences use the recursive
        // residual). Restarts stop when one no longer halves the true
        // residual (the FP32 iterate's attainable floor); the status then
        // stays the recursive criterion's, and last_true_residual() says how
        // far the true residual got.
        for (;;) {
            if (!nu_true) {
                check(launch_residual64_absmax(v0, x0_, b0_, r0_, sc_, partials_, mailbox_dev_, stream_),
                      "residual64");
                if (wait_nu(&nu)) return -3;
            }
            true_nu_ = nu;
            if (nu <= numax || !(nu < 0.5f * prev)) break;
            if (restarts >= kMaxReplacements || it >= opt_.max_iterations) break;
            prev = nu;
            ++restarts;
            ++replacements_;
            status = pcg_iterate(nu, numax, it, nu_true, restart_below(), false);
            if (status < 0) return status;
            if (status != kSuccess) break;
        }
    }
    if (iterations) *iterations = it;
    check(launch_gather_dof(v0, x0_, x_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "solve sync")) return -3;
    return first_error_ == hipSuccess ? status : -3;
}

// Lines 6-16 of solveMultigridPCG from the residual in r0_ (nu = its
// ||r||_inf): p = M r, then the iteration pieces until ||r||_inf <= numax
// (x updated, kSuccess), the residual grows after this recurrence's third
// iteration or `it` reaches max_iterations (kFailed), or an error (< 0).
// With `replace`, a piece started while ||r||_inf <= kReplaceBand * numax
// forms r as the FP64 true residual b - A x of the new iterate instead of
// r - alpha A p (counted in replacements_). When such a true residual does
// not halve, the recurrence ends (x updated, kSuccess with nu above numax)
// if nu < restart_below, so the caller restarts it; otherwise it goes on
// with the recursive residual. nu_true says that the final nu is the FP64
// true residual of the returned x.
int PoissonDevice::pcg_iterate(float& nu, float numax, int& it, bool& nu_true, float restart_below, bool replace) {
    const LevelView& v0 = lv_[0].v;
    const int it0 = it;   // the growth test counts this recurrence's own iterations
    nu_true = false;
    if (run(g_first_, [&] { issue_first(); })) return -3;
    int k = 0;   // p_[k] holds the current direction
    bool replaced = false, stalled = false;
    for (;;) {
        const float nu_old = nu;
        if (wait_nu(&nu)) return -3;
        if (nu <= numax) {
            check(launch_update_x(v0, p_[k], x0_, sc_, stream_), "update x");
            nu_true = replaced;
            return kSuccess;
        }
        if (replaced && !(nu < 0.5f * nu_old)) {
            // the true residual stopped halving: a fresh recurrence from it
            // when the restart rule allows one, else this one goes on with
            // the recursive residual
            if (nu < restart_below) {
                check(launch_update_x(v0, p_[k], x0_, sc_, stream_), "update x");
                nu_true = true;
                return kSuccess;
            }
            stalled = true;
        } else if (nu > nu_old && it - it0 > 3) {
            return kFailed;
        }
        replaced = replace && !stalled && nu <= kReplaceBand * numax;
        if (replaced) {
            ++replacements_;
            if (run(g_repl_[k], [&] { issue_iter(k, true); })) return -3;
        } else if (run(g_iter_[k], [&] { issue_iter(k); })) {
            return -3;
        }
        k ^= 1;
        if (++it >= opt_.max_iterations) {
            if (wait_nu(&nu)) return -3;   // the piece's own reduction
            return kFailed;
        }
    }
}

// solvePureMultigrid (simd_vdb_poisson.cpp:2228-2267); issued directly (the
// fallback path after a FAILED PCG is not latency critical)
int PoissonDevice::solve_mg_dev(float* x_dof, const float* b_dof, int* iterations) {
    if (!on_device()) return -4;
    if (first_error_ != hipSuccess) return -3;
    if (follow_caller()) return -3;
    if (opt_.smoother == kSpai0 && ensure_spai0()) return -3;
    const LevelView& v0 = lv_[0].v;
    int it = 0;
    if (iterations) *iterations = 0;
    check(launch_scatter_dof(v0, x_dof, x0_, stream_), "scatter x");
    check(launch_scatter_dof(v0, b_dof, b0_, stream_), "scatter b");
    check(launch_apply(v0, kModeResidual, x0_, b0_, r0_, 0.f, stream_), "residual");
    check(launch_absmax(v0, r0_, sc_, partials_, mailbox_dev_, stream_), "absmax");
    float nu;
    if (wait_nu(&nu)) return -3;
    const float numax = opt_.relative_tolerance * nu;
    int status = kFailed;
    if (nu <= numax) {
        status = kSuccess;
    } else {
        for (; it < opt_.max_iterations; ++it) {
            mg_cycle(0, x0_, b0_);
            check(launch_apply(v0, kModeResidual, x0_, b0_, r0_, 0.f, stream_), "residual");
            check(launch_absmax(v0, r0_, sc_, partials_, mailbox_dev_, stream_), "absmax");
            const float nu_old = nu;
            if (wait_nu(&nu)) return -3;
            if (nu <= numax) {
                status = kSuccess;
                break;
            }
            if (nu > nu_old && it > 8) {
                status = kFailed;
                break;
            }
        }
    }
    if (iterations) *iterations = it;
    check(launch_gather_dof(v0, x0_, x_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "solve sync")) return -3;
    return first_error_ == hipSuccess ? status : -3;
}

int PoissonDevice::apply_dev(int l, int mode, float w, const float* x_dof, const float* b_dof, float* y_dof) {
    if (l < 0 || l >= nlevels()) return -1;
    if (mode == kModeSpai0 && ensure_spai0()) return -3;
    const LevelView& v = lv_[l].v;
    float *xb = staging_[0], *bb = staging_[1], *yb = staging_[2];
    check(launch_scatter_dof(v, x_dof, xb, stream_), "scatter");
    if (b_dof) check(launch_scatter_dof(v, b_dof, bb, stream_), "scatter");
    float* out = yb;
    if (mode == kModeRedGS || mode == kModeBlackGS) out = xb;   // in place
    check(launch_apply(v, mode, xb, bb, out, w, stream_), "apply");
    check(launch_gather_dof(v, out, y_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "sync")) return -3;
    return 0;
}

int PoissonDevice::first_sweep_dev(int l, int kind, float w, const float* b_dof, float* y_dof) {
    if (l < 0 || l >= nlevels() || kind < 0 || kind > 2) return -1;
    if (kind == 1 && ensure_spai0()) return -3;
    const LevelView& v = lv_[l].v;
    float *bb = staging_[1], *yb = staging_[2];
    check(launch_scatter_dof(v, b_dof, bb, stream_), "scatter");
    // the staging bricks are shared by all levels: rows without DOFs must read
    // as zero for the black pass that follows the red update
    check(hipMemsetAsync(yb, 0, size_t(v.nleaf) * kLeafSize * sizeof(float), stream_), "memset");
    check(launch_first_sweep(v, kind, bb, yb, w, stream_), "first");
    if (kind == 2) check(launch_apply(v, kModeBlackGS, yb, bb, yb, w, stream_), "black");
    check(launch_gather_dof(v, yb, y_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "sync")) return -3;
    return 0;
}

int PoissonDevice::restrict_dev(int fl, const float* fine_dof, float* coarse_dof) {
    if (fl < 0 || fl + 1 >= nlevels()) return -1;
    check(launch_scatter_dof(lv_[fl].v, fine_dof, staging_[0], stream_), "scatter");
    check(launch_restrict(lv_[fl + 1].v, staging_[0], staging_[1], stream_), "restrict");
    check(launch_gather_dof(lv_[fl + 1].v, staging_[1], coarse_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "sync")) return -3;
    return 0;
}

int PoissonDevice::prolong_dev(int fl, const float* coarse_dof, float* fine_dof, float alpha) {
    if (fl < 0 || fl + 1 >= nlevels()) return -1;
    check(launch_scatter_dof(lv_[fl + 1].v, coarse_dof, staging_[1], stream_), "scatter");
    check(launch_scatter_dof(lv_[fl].v, fine_dof, staging_[0], stream_), "scatter");
    check(launch_prolong_add(lv_[fl].v, staging_[1], staging_[0], alpha, stream_), "prolong");
    check(launch_gather_dof(lv_[fl].v, staging_[0], fine_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "sync")) return -3;
    return 0;
}

int PoissonDevice::precondition_dev(const float* r_dof, float* z_dof) {
    if (opt_.smoother == kSpai0 && ensure_spai0()) return -3;
    const LevelView& v0 = lv_[0].v;
    check(launch_scatter_dof(v0, r_dof, r0_, stream_), "scatter");
    precond(0, z0_, r0_, true, -1);
    check(launch_gather_dof(v0, z0_, z_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "sync")) return -3;
    return 0;
}

double PoissonDevice::algorithmic_bytes(int l, int mode) const {
    if (l < 0 || l >= nlevels()) return 0.0;
    const PoissonLevelHost& h = host_.levels[l];
    const double sv = (mixed_ && l > 0) ? 2.0 : 4.0;   // bytes of a W-cycle level vector entry
    if (mode == kModeCoarsest) {
        // b in, x out, and the system's rows (6 neighbour indices, 8 coefficients, 1 position) once
        return double(h.ndof) * (2.0 * sv + 24.0 + 32.0 + 4.0);
    }
    double total = 0.0;
    if (mode == kModeBlockPost && l + 1 < nlevels()) total += sv * double(host_.levels[l + 1].ndof);   // coarse x
    if (mode == kModeBlockPost && post_cg_level(l)) total += algorithmic_bytes(l + 1, kModeCoarsest);   // its solve
    for (size_t i = 0; i < h.nleaf(); ++i) {
        const double nd = mask_count(&h.mask[i * 8]);
        const int kd = h.keep[kCoefDiag][i], kx = h.keep[kCoefX][i], ky = h.keep[kCoefY][i], kz = h.keep[kCoefZ][i];
        const int ks = h.keep[kCoefSpai0].empty() ? 0 : h.keep[kCoefSpai0][i];
        const double c = 4.0 * (kd + kx + ky + kz), d = 4.0 * kd;
        double per = 0;
        switch (mode) {
            case kModeLaplacian: per = 8 + c; break;
            case kModeResidual: per = 12 + c; break;
            case kModeJacobi: per = 12 + c + d; break;
            case kModeSpai0: per = 12 + c + 4.0 * ks; break;
            case kModeJacobiBf16: per = (l == 0 ? 8 : 6) + c + d; break;   // x, y at 2 B; b at 4 B on level 0
            // fused prolongation + three sweeps: x in, b, x out, the coefficients once
            case kModeBlockPost: per = 3.0 * sv + c + d; break;
            default: per = 12 + 4.0 * (kx + ky + kz) + d; break;   // one GS colour pass
        }
        total += nd * per;
    }
    return total;
}

int PoissonDevice::time_apply(int l, int mode, int reps, float* ms) {
    if (!on_device() || l < 0 || l >= nlevels() || reps <= 0 || mode < 0 || mode > kModeCoarsest) return -1;
    if (mode == kModeJacobiBf16 && !mixed_) return -1;
    if (mode == kModeBlockPost && (l == 0 || l + 1 >= nlevels() || !lv_[l].v.nb27)) return -1;
    if (mode == kModeCoarsest && l != nlevels() - 1) return -1;
    if (mode == kModeSpai0 && ensure_spai0()) return -3;
    const LevelView& v = lv_[l].v;
    const float* X = l == 0 ? x0_ : lv_[l].X;
    const float* B = l == 0 ? b0_ : lv_[l].B;
    float* Y = staging_[2];
    DevLevel& d = lv_[l];
    // the BF16 sweep as the mixed W-cycle runs it: level 0 reads the FP32 residual; the block post-visit
    // and the coarsest CG on the level's own scratch vectors (every solve overwrites them before use)
    auto launch = [&]() {
        if (mode == kModeBlockPost) {
            DevLevel& c = lv_[l + 1];
            const float w0 = scheduled_weight(kSweeps, kSweeps - 1), w1 = scheduled_weight(kSweeps, kSweeps - 2),
                        w2 = scheduled_weight(kSweeps, kSweeps - 3);
            if (post_cg_level(l)) {
                if (mixed_)
                    return launch_block_post_cg(v, cv_, coarsest_dof_of_, c.Bh, c.Xh, d.Th, d.Bh, d.Xh, w0, w1, w2,
                                                stream_);
                return launch_block_post_cg(v, cv_, coarsest_dof_of_, c.B, c.X, d.T, d.B, d.X, w0, w1, w2, stream_);
            }
            if (mixed_) return launch_block_post(v, d.Th, c.Xh, d.Bh, d.Xh, w0, w1, w2, stream_);
            return launch_block_post(v, d.T, c.X, d.B, d.X, w0, w1, w2, stream_);
        }
        if (mode == kModeCoarsest) {
            if (mixed_) return launch_coarsest_cg(cv_, d.Bh, d.Xh, stream_);
            return launch_coarsest_cg(cv_, d.B, d.X, stream_);
        }
        if (mode != kModeJacobiBf16) return launch_apply(v, mode, X, B, Y, 0.8571f, stream_);
        if (l == 0) return launch_jacobi(v, d.Th, b0_, d.Uh, 0.8571f, stream_);
        return launch_jacobi(v, d.Th, d.Bh, d.Uh, 0.8571f, stream_);
    };
    hipEvent_t e0, e1;
    if (check(hipEventCreate(&e0), "event") || check(hipEventCreate(&e1), "event")) return -3;
    check(launch(), "warm");
    check(hipEventRecord(e0, stream_), "event");
    for (int i = 0; i < reps; ++i) check(launch(), "apply");
    check(hipEventRecord(e1, stream_), "event");
    check(hipEventSynchronize(e1), "event sync");
    float total = 0.f;
    check(hipEventElapsedTime(&total, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (first_error_ != hipSuccess) return -3;
    *ms = total / float(reps);
    return 0;
}

int PoissonDevice::coarsest_dev(const float* b_dof, float* x_dof) {
    const int l = nlevels() - 1;
    const LevelView& v = lv_[l].v;
    check(launch_scatter_dof(v, b_dof, staging_[1], stream_), "scatter");
    check(hipMemsetAsync(staging_[2], 0, size_t(v.nleaf) * kLeafSize * sizeof(float), stream_), "memset");
    check(launch_coarsest_cg(cv_, staging_[1], staging_[2], stream_), "coarsest");
    check(launch_gather_dof(v, staging_[2], x_dof, stream_), "gather");
    if (check(hipStreamSynchronize(stream_), "sync")) return -3;
    return 0;
}

}  // namespace uaamg
