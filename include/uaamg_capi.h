/*
 * uaamg_capi.h - C ABI of the MI355X UAAMG-PCG pressure-Poisson solver.
 *
 * The reference exposes the solve path as C++ classes in namespace
 * simd_uaamg that take OpenVDB grids (SIMD_UAAMG/simd_vdb_poisson.h). This
 * header replaces that surface with plain pointers and sizes: an OpenVDB grid
 * is passed as its flattened leaves (origins, active masks, leaf buffers), and
 * vectors are passed DOF-ordered, i.e. in the order LaplacianWithLevel::
 * gridToVector produces (simd_vdb_poisson.cpp:1584-1600). Each entry point
 * cites the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns an int status: 0 on success, 1 = FAILED for the
 *    solve calls (PoissonSolver::FAILED, simd_vdb_poisson.h:151), negative on
 *    error (see UAAMG_E*). uaamg_last_error() describes the last error of the
 *    calling thread. No C++ exception crosses this boundary.
 *  - Host arrays are borrowed for the duration of a call only.
 *  - A solver handle owns its device memory; use one handle per host thread.
 */
#ifndef UAAMG_CAPI_H
#define UAAMG_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UAAMG_SUCCESS 0
#define UAAMG_FAILED 1
#define UAAMG_EINVAL (-1)
#define UAAMG_ENOMEM (-2)
#define UAAMG_EHIP (-3)
#define UAAMG_ENODEV (-4)

/* PoissonSolver::SmootherOption (simd_vdb_poisson.h:143-148) */
#define UAAMG_SMOOTHER_SCHEDULED_RELAXED_JACOBI 0
#define UAAMG_SMOOTHER_WEIGHTED_JACOBI 1
#define UAAMG_SMOOTHER_RED_BLACK_GAUSS_SEIDEL 2
#define UAAMG_SMOOTHER_SPAI0 3

/* uaamg_options.precision */
#define UAAMG_PRECISION_FP32 0       /* FP32 throughout (the reference's arithmetic) */
#define UAAMG_PRECISION_BF16_CYCLE 1 /* W-cycle level vectors stored in BF16, computed in FP32, for the
                                        Jacobi-family smoothers (RBGS / SPAI0 cycles stay FP32); outer
                                        PCG in FP32 with FP64 dot products and an FP64 residual
                                        replacement after convergence. No reference counterpart. */

/* LaplacianApplySIMD::WorkingMode (simd_vdb_poisson.cpp:683-690) */
#define UAAMG_MODE_LAPLACIAN 0
#define UAAMG_MODE_RESIDUAL 1
#define UAAMG_MODE_WEIGHTED_JACOBI 2
#define UAAMG_MODE_SPAI0 3
#define UAAMG_MODE_RED_GS 4
#define UAAMG_MODE_BLACK_GS 5
#define UAAMG_MODE_JACOBI_BF16 6 /* algorithmic_bytes / time_apply only: the BF16 W-cycle's Jacobi sweep */
#define UAAMG_MODE_BLOCK_POST 7  /* algorithmic_bytes / time_apply only: a coarse level's post-visit block kernel
                                    (prolongation and three sweeps in one launch; levels with block tables) */
#define UAAMG_MODE_COARSEST 8    /* algorithmic_bytes / time_apply only: the coarsest-level CG (its level) */

/* coefficient kinds of a level */
#define UAAMG_COEF_DIAGONAL 0
#define UAAMG_COEF_X 1
#define UAAMG_COEF_Y 2
#define UAAMG_COEF_Z 3
#define UAAMG_COEF_SPAI0 4

/*
 * A flattened OpenVDB FloatGrid (ncomp = 1) or Vec3fGrid (ncomp = 3).
 * origins[i*3..i*3+2] is the origin of leaf i (a multiple of 8); leaves may come
 * in any order and are put in OpenVDB Tree4<5,4,3> traversal order internally.
 * values holds each leaf's buffer: 512 entries (512*3 for Vec3f, AoS) at voxel
 * offset 64x+8y+z. active[i*8+w] bit b is the active state of offset 64w+b
 * (NULL = all active). Lookups outside the leaves return background.
 */
typedef struct uaamg_grid {
    int64_t nleaf;
    const int32_t* origins;
    const uint64_t* active;
    const float* values;
    int32_t ncomp;
    float background[3];
} uaamg_grid;

/* PoissonSolver public fields (simd_vdb_poisson.h:165-169) plus device placement */
typedef struct uaamg_options {
    float relative_tolerance; /* mRelativeTolerance, default 1e-7 */
    int32_t max_iterations;   /* mMaxIteration, default 100 */
    int32_t smoother;         /* mSmoother, UAAMG_SMOOTHER_* */
    int32_t precision;        /* UAAMG_PRECISION_*, default FP32 */
    int32_t device;           /* HIP device ordinal */
    int32_t max_coarsest_dof; /* coarsen while ndof > this (4000, simd_vdb_poisson.cpp:1815) */
    int32_t flags;            /* 0 normally; bit 0 = unfused kernels, bit 1 = no graph replay,
                                 bit 2 = bottom levels in one single-workgroup launch, bit 4 = no
                                 block-visit kernels (A/B checks); bit 6 = FP64 residual replacement
                                 at FP32 precision; bit 7 = pure-multigrid fallback in the merged batch
                                 and the solve server: a problem whose PCG returns FAILED is zeroed and
                                 solved by solvePureMultigrid, as FLIP_vdb.cpp:3150-3155 does; bit 8 =
                                 the whole hierarchy build on the host threads instead of the device
                                 (same result: by default uaamg_poisson_create builds every level on
                                 the device and keeps the coefficient bricks there) */
    void* stream;             /* hipStream_t to run on; NULL = a stream owned by the handle, which each device
                                 entry point first orders after the work queued on the null stream */
} uaamg_options;

typedef struct uaamg_poisson uaamg_poisson;

/* Fills the reference defaults (PoissonSolver ctor, simd_vdb_poisson.h:153-160). */
int uaamg_options_default(uaamg_options* opts);

/* Text of the calling thread's last error ("" if none). */
const char* uaamg_last_error(void);

/* Library version string. */
const char* uaamg_version(void);

/*
 * Builds the level-0 Laplacian and the aggregation hierarchy and uploads it.
 * Replaces LaplacianWithLevel::createPressurePoissonLaplacian (simd_vdb_poisson.h:54-57,
 * .cpp:95-99) followed by PoissonSolver(LaplacianWithLevel::Ptr) (simd_vdb_poisson.h:153-160).
 * dx is liquid_phi->voxelSize()[0]. A liquid phi without degrees of freedom (no active voxel with
 * phi < 0) is rejected with UAAMG_EINVAL: the reference's solve of that empty system returns
 * SUCCESS after 0 iterations and changes nothing, so a caller with no liquid skips the solve.
 */
int uaamg_poisson_create(const uaamg_grid* liquid_phi, const uaamg_grid* face_weights, float dt, float dx,
                         const uaamg_options* opts, uaamg_poisson** out);

void uaamg_poisson_destroy(uaamg_poisson* h);

/* Sets mRelativeTolerance / mMaxIteration / mSmoother (simd_vdb_poisson.h:165-169). */
int uaamg_poisson_set_options(uaamg_poisson* h, const uaamg_options* opts);

/* mNumDof of level 0 (simd_vdb_poisson.h:124). */
int64_t uaamg_poisson_ndof(const uaamg_poisson* h);

/* mMultigridHierarchy.size() (simd_vdb_poisson.h:170). */
int uaamg_poisson_nlevels(const uaamg_poisson* h);

/*
 * Per-level sizes: DOF count, leaves holding DOFs, and leaves of the reference's
 * DOF tree (which also keeps empty leaves).
 */
int uaamg_poisson_level_info(const uaamg_poisson* h, int level, int64_t* ndof, int64_t* nleaf_dof,
                             int64_t* nleaf_tree);

/*
 * DOF topology of a level (mDofIndex, simd_vdb_poisson.h:121): origins [nleaf_dof][3], masks [nleaf_dof][8]
 * and the index of each leaf's first DOF [nleaf_dof]. Any output may be NULL.
 * Implements the data behind getZeroVectorGrid / the mDofLeafManager visitor.
 */
int uaamg_poisson_level_topology(const uaamg_poisson* h, int level, int32_t* origins, uint64_t* masks,
                                 int64_t* dofbase);

/*
 * Effective coefficient bricks [nleaf_dof][512] of a level (mDiagonal, mXEntry, mYEntry, mZEntry,
 * simd_vdb_poisson.h:116-119; SPAI0 of LaplacianApplySIMD::init_SPAI0) as an accessor reads them,
 * and kept[i] = 0 where the reference trims the leaf to the default (trimDefaultNodes, :1756-1811).
 */
int uaamg_poisson_level_coefficients(const uaamg_poisson* h, int level, int kind, float* bricks, uint8_t* kept);

/*
 * LaplacianWithLevel::gridToVector (simd_vdb_poisson.cpp:1584-1600): vec[d] = the grid's stored value
 * (active or not) at DOF d's voxel, where the grid has a leaf at that DOF leaf's origin, else 0.
 * vec: ndof floats, DOF order.
 */
int uaamg_poisson_grid_to_vector(const uaamg_poisson* h, const uaamg_grid* grid, float* vec);

/*
 * LaplacianWithLevel::vectorToGrid (simd_vdb_poisson.cpp:1602-1621) onto the DOF topology of
 * uaamg_poisson_level_topology(level 0): bricks [nleaf_dof][512] receive vec at the DOF voxels (which
 * the grid sets active) and 0 elsewhere (the topology copy's background). Used by getZeroVectorGrid
 * and by the C++ facade to write a solve's result into the caller's grid.
 */
int uaamg_poisson_vector_to_grid(const uaamg_poisson* h, const float* vec, float* bricks);

/*
 * LaplacianWithLevel::getTriplets (simd_vdb_poisson.cpp:129-206) of a level: *count receives the
 * number of triplets; when rows / cols / vals are non-NULL they receive them (row DOF, column DOF,
 * value) in the reference's order - DOFs ascending, per DOF the diagonal, then per axis the lower
 * and the upper neighbour where that neighbour is a DOF.
 */
int uaamg_poisson_level_triplets(const uaamg_poisson* h, int level, int64_t* count, int32_t* rows, int32_t* cols,
                                 float* vals);

/*
 * b = -(1/dx) sum_faces +-(w u + (1-w) u_solid), DOF-ordered (ndof floats).
 * Replaces LaplacianWithLevel::createPressurePoissonRightHandSide (simd_vdb_poisson.h:59-65, .cpp:101-127).
 * solid_velocity may be NULL (zero).
 */
int uaamg_poisson_build_rhs(const uaamg_poisson* h, const uaamg_grid* face_weights, const uaamg_grid* vx,
                            const uaamg_grid* vy, const uaamg_grid* vz, const uaamg_grid* solid_velocity,
                            float* rhs);

/*
 * The same right-hand side built on the device into rhs_dev (ndof floats, DOF order, device memory),
 * bit-identical to uaamg_poisson_build_rhs: one wavefront per level-0 DOF leaf reads the face weights,
 * liquid and solid velocities at each DOF voxel and at its +x/+y/+z neighbour voxel. The grids' leaves
 * are matched to the DOF leaves by origin on the host (leaf tables, O(leaves)); values_on_device = 0:
 * each grid's values are host memory, uploaded for the call; 1: they are device memory
 * ([nleaf][512][ncomp] in the grid's own leaf order; origins stay host memory). solid_velocity may be
 * NULL (zero). Allocates device memory for the call (values when uploaded, the leaf tables) and
 * returns after rhs_dev is written.
 * Replaces LaplacianWithLevel::createPressurePoissonRightHandSide (simd_vdb_poisson.h:59-65, .cpp:101-127)
 * for callers whose velocities live on the GPU, or whose host cannot afford the per-voxel loop.
 */
int uaamg_poisson_build_rhs_device(uaamg_poisson* h, const uaamg_grid* face_weights, const uaamg_grid* vx,
                                   const uaamg_grid* vy, const uaamg_grid* vz, const uaamg_grid* solid_velocity,
                                   int32_t values_on_device, float* rhs_dev);

/* Level-0 DOF leaves uaamg_poisson_build_rhs_device_noalloc takes at most. */
#define UAAMG_RHS_NOALLOC_MAX_LEAVES 131072

/*
 * uaamg_poisson_build_rhs_device with no device allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory - code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc. Every grid's values must already be device memory (as with
 * values_on_device = 1); the leaf tables go into a fixed scratch array of the library (10 MiB per
 * GPU, held until the process ends) instead of a per-call allocation. Returns UAAMG_EINVAL when
 * level 0 has more than UAAMG_RHS_NOALLOC_MAX_LEAVES DOF leaves. All calls on one GPU share that
 * array: calls that overlap on one GPU (several host threads) race on it and must be serialised by
 * the caller. Same result as uaamg_poisson_build_rhs_device, bit for bit.
 */
int uaamg_poisson_build_rhs_device_noalloc(uaamg_poisson* h, const uaamg_grid* face_weights, const uaamg_grid* vx,
                                           const uaamg_grid* vy, const uaamg_grid* vz,
                                           const uaamg_grid* solid_velocity, float* rhs_dev);

/*
 * One face-velocity channel of the caller's grid, updated in place: leaves in any order, active
 * masks [nleaf][8] (host memory, in/out), values [nleaf][512] (in/out) and the optional outputs
 * update / reflected [nleaf][512] (NULL: not produced). values, update and reflected are host or
 * device memory as the call's values_on_device says.
 */
typedef struct uaamg_face_channel {
    int64_t nleaf;
    const int32_t* origins;
    uint64_t* active;
    float* values;
    float* update;
    float* reflected;
} uaamg_face_channel;

/*
 * The step after the solve on the device: the pressure gradient applied to the face velocities.
 * For every active face of channel c with weight w > 0 where the pressure has a DOF on either side,
 *   du = -dt (p(v) - p(v - e_c)) / dx / theta, theta = fraction_inside(phi(v - e_c), phi(v))
 *   clamped to >= 0.02 when either phi >= 0 (else 1); u += du; reflected = u + du; and where w < 1
 *   u = (1 - s) u + s u_solid with s = (1 - w) friction;
 * every other active face is turned off. update receives du on updated faces and the incoming
 * velocity elsewhere (the reference's deep copy), reflected the reflected velocity likewise; the
 * active masks end as the updated faces. pressure_dev: the solve's DOF-ordered result (device).
 * Replaces FLIP_Solver_OpenVDB::apply_pressure_gradient (FLIP_vdb.cpp:2571-2667); bit-identical to
 * its restatement. channels: 3 entries (x, y, z); solid_velocity may be NULL (zero).
 */
int uaamg_poisson_apply_pressure_gradient_device(uaamg_poisson* h, const float* pressure_dev,
                                                 const uaamg_grid* liquid_phi, const uaamg_grid* face_weights,
                                                 const uaamg_grid* solid_velocity, uaamg_face_channel* channels,
                                                 float dt, float friction, int32_t values_on_device);

/*
 * The face weights of the pressure solve from the solid SDF, computed on HIP device `device`.
 * The weight grid has the liquid SDF's topology dilated twice over the 26 neighbours; at each active
 * voxel v the weight of face c is 1 - fraction_inside(four solid SDF values at the corners v + {0,1}^2
 * of that face's plane), clamped to [0, 1]; every other voxel of the grid holds 1 (the background).
 * Two-phase: *nleaf receives the leaf count; when origins ([nleaf][3]), active ([nleaf][8]) and values
 * ([nleaf][512][3], host memory) are non-NULL they receive the grid, leaves in OpenVDB tree order.
 * Replaces FLIP_Solver_OpenVDB::calculate_face_weights (FLIP_vdb.cpp:1990-2068).
 */
int uaamg_face_weights(const uaamg_grid* liquid_sdf, const uaamg_grid* solid_sdf, int32_t device, int64_t* nleaf,
                       int32_t* origins, uint64_t* active, float* values);

/* variants of uaamg_face_weights_ex */
#define UAAMG_FACE_WEIGHTS_STATIC 0 /* FLIP_Solver_OpenVDB::calculate_face_weights(liquid, solid), FLIP_vdb.cpp:1990-2068 */
#define UAAMG_FACE_WEIGHTS_MEMBER 1 /* the member calculate_face_weights() FLIP's substep calls, FLIP_vdb.cpp:2669-2752:
                                       the same weights, each one below 1e-2 after the clamp set to 0 */

/* uaamg_face_weights with the variant chosen (uaamg_face_weights = variant UAAMG_FACE_WEIGHTS_STATIC). */
int uaamg_face_weights_ex(const uaamg_grid* liquid_sdf, const uaamg_grid* solid_sdf, int32_t device, int32_t variant,
                          int64_t* nleaf, int32_t* origins, uint64_t* active, float* values);

/*
 * PoissonSolver::solveMultigridPCG (simd_vdb_poisson.h:162, .cpp:2155-2226) on host DOF-ordered
 * vectors: x holds the initial guess on entry and the pressure on exit. Returns UAAMG_SUCCESS or
 * UAAMG_FAILED; *iterations receives mIterationTaken.
 */
int uaamg_poisson_solve_pcg(uaamg_poisson* h, float* x, const float* b, int32_t* iterations);

/* Same with x and b in device memory (DOF-ordered, ndof floats each). */
int uaamg_poisson_solve_pcg_device(uaamg_poisson* h, float* x_dev, const float* b_dev, int32_t* iterations);

/*
 * After a solve with FP64 residual replacement (UAAMG_PRECISION_BF16_CYCLE, or flags bit 6):
 * ||b - A x||_inf of the returned x evaluated in FP64, and the number of times the PCG residual
 * was replaced by that true residual: the first recurrence's steps taken within 1000x of the
 * tolerance, plus restarts of the recurrence when the true residual missed the tolerance (after
 * the recursive one met it, or when a replaced step stopped halving it).
 * Both are -1 when the last solve ran without replacement. No reference counterpart.
 */
int uaamg_poisson_last_residual(const uaamg_poisson* h, float* true_residual_inf, int32_t* replacements);

/*
 * Many independent solves at once (batch configuration: one handle per problem, each with its own
 * stream). Up to `concurrency` (<= 16) host threads each drive one handle's solve loop, so small
 * problems overlap on the GPU. status[i] receives that problem's UAAMG_SUCCESS / UAAMG_FAILED,
 * iterations[i] its mIterationTaken (may be NULL). No reference counterpart: FLIP solves one
 * problem per call (FLIP_vdb.cpp:3132-3157); the batch is this solver's entry point for callers that
 * hold many problems. A handle may appear only once per call.
 */
int uaamg_poisson_solve_batch_device(uaamg_poisson* const* handles, int32_t n, float* const* x_dev,
                                     const float* const* b_dev, int32_t* status, int32_t* iterations,
                                     int32_t concurrency);

/*
 * The merged batch (batch configuration, one launch per W-cycle stage): n independent problems whose
 * level-l leaves are merged into one leaf table per level (deepest hierarchies first, per-problem
 * coefficient defaults, level-0 ranges aligned to the reduction block), solved together: every
 * smoothing sweep, transfer, coarsest group and PCG vector operation is one launch for the whole
 * batch, one hipGraph covers one PCG iteration of all problems, and each problem keeps its own PCG
 * scalars and convergence state on the device. Each problem's iterates equal a solve of it alone
 * (same kernels, same reduction partition and order). opts applies to every problem (Jacobi-family
 * smoothers; precision FP32 or BF16 W-cycle; flags bit 1 = no graph replay, bit 4 = no block-visit
 * kernels). No reference counterpart (FLIP solves one problem per call, FLIP_vdb.cpp:3132-3157).
 */
typedef struct uaamg_poisson_batch uaamg_poisson_batch;
int uaamg_poisson_batch_create(int32_t n, const uaamg_grid* const* liquid_phi, const uaamg_grid* const* face_weights,
                               const float* dt, const float* dx, const uaamg_options* opts,
                               uaamg_poisson_batch** out);
void uaamg_poisson_batch_destroy(uaamg_poisson_batch* h);
/* mRelativeTolerance / mMaxIteration for the next solves (smoother and precision are fixed). */
int uaamg_poisson_batch_set_options(uaamg_poisson_batch* h, const uaamg_options* opts);
/* problems, merged levels, merged level-0 leaves (with alignment padding), device bytes */
int uaamg_poisson_batch_info(const uaamg_poisson_batch* h, int32_t* nproblems, int32_t* nlevels, int64_t* leaves0,
                             int64_t* device_bytes);
/* mNumDof of problem i (-1 for a bad index) */
int64_t uaamg_poisson_batch_ndof(const uaamg_poisson_batch* h, int32_t i);
/* Solves every problem: x_dev[i] / b_dev[i] are DOF-ordered device vectors of problem i (x: initial
 * guess in, solution out); status[i] = UAAMG_SUCCESS / UAAMG_FAILED, iterations[i] its
 * mIterationTaken (either may be NULL). Returns 0 or a negative error. */
int uaamg_poisson_batch_solve_device(uaamg_poisson_batch* h, float* const* x_dev, const float* const* b_dev,
                                     int32_t* status, int32_t* iterations);
/* Per problem after the last solve, as uaamg_poisson_last_residual (arrays of nproblems entries). */
int uaamg_poisson_batch_last_residual(const uaamg_poisson_batch* h, float* true_residual_inf, int32_t* replacements);
/* After a solve with flags bit 7 (FLIP's fallback, FLIP_vdb.cpp:3150-3155): per input problem, the
 * status and iteration count (mIterationTaken) of its solvePureMultigrid run, -1 where its PCG
 * succeeded. The PCG's FAILED status and iterations stay in uaamg_poisson_batch_solve_device's
 * outputs; x holds the pure-multigrid result. */
int uaamg_poisson_batch_last_fallback(const uaamg_poisson_batch* h, int32_t* mg_status, int32_t* mg_iterations);

/* PoissonSolver::solvePureMultigrid (simd_vdb_poisson.h:163, .cpp:2228-2267), host vectors. */
int uaamg_poisson_solve_mg(uaamg_poisson* h, float* x, const float* b, int32_t* iterations);

/*
 * One operator sweep on a level with host DOF-ordered vectors (LaplacianWithLevel::laplacianApply /
 * residualApply / weightedJacobiApply / spai0Apply, simd_vdb_poisson.cpp:1347-1385). For
 * UAAMG_MODE_RED_GS / _BLACK_GS this is one colour pass in place: y = pass(x), weight = SOR weight.
 * b may be NULL for UAAMG_MODE_LAPLACIAN.
 */
int uaamg_poisson_apply(uaamg_poisson* h, int level, int mode, float weight, const float* x, const float* b,
                        float* y);

/*
 * First smoothing sweep from a zero guess (setGridToResultAfterFirstJacobi / SPAI0 / RBGS<true>,
 * simd_vdb_poisson.cpp:1435-1582): kind 0 = Jacobi with weight, 1 = SPAI0, 2 = red-black GS with SOR weight.
 */
int uaamg_poisson_first_sweep(uaamg_poisson* h, int level, int kind, float weight, const float* b, float* y);

/* LaplacianWithLevel::restriction from fine_level to fine_level+1 (simd_vdb_poisson.cpp:1624-1684). */
int uaamg_poisson_restrict(uaamg_poisson* h, int fine_level, const float* fine, float* coarse);

/* LaplacianWithLevel::prolongation<true>: fine += alpha * coarse (simd_vdb_poisson.cpp:1686-1754). */
int uaamg_poisson_prolong(uaamg_poisson* h, int fine_level, const float* coarse, float* fine, float alpha);

/* One application of the W-cycle preconditioner z = M r (muCyclePreconditioner<2,true>, :1844-1977). */
int uaamg_poisson_precondition(uaamg_poisson* h, const float* r, float* z);

/* The coarsest-level solve (Eigen CG, 10 iterations from zero; :2114-2152, 1870-1875). */
int uaamg_poisson_coarsest_solve(uaamg_poisson* h, const float* b, float* x);

/*
 * Measurement hooks (no reference counterpart; the reference only times whole solves with
 * CSim::TimerMan). algorithmic_bytes: bytes one sweep of `mode` on `level` must move (each stream
 * once per DOF; coefficient streams only for leaves that keep stored coefficients).
 * time_apply: mean milliseconds per launch of `reps` back-to-back sweeps on the handle's stream,
 * measured with HIP events on that stream.
 */
int uaamg_poisson_algorithmic_bytes(const uaamg_poisson* h, int level, int mode, double* bytes);
/* Level-to-kernel mapping chosen when the handle was created: *block = 1 when a W-cycle visit of the
 * level runs as two block-visit launches (2: and the second of them also runs the coarsest solve),
 * and the ms per visit pair of both mappings as timed then (-1 when the level was not timed).
 * No reference counterpart. */
int uaamg_poisson_kernel_map(const uaamg_poisson* h, int level, int32_t* block, float* ms_block, float* ms_sweep);
int uaamg_poisson_time_apply(uaamg_poisson* h, int level, int mode, int reps, float* ms_per_launch);

/* ---------------------------------------------------------------- scenes */
/*
 * Synthetic scenes on [0,1]^3 (liquid SDF, face weights, seeded face velocities) used to drive
 * the solver without a simulator: config 0 = pool 40% + sphere r=0.15; 1 = pool 30% + 8 seeded
 * spheres + solid sphere r=0.2 at the centre; 3 = randomised batch problem. n <= 0: preset edge.
 * These stand in for the caller's grids (FLIP_vdb.cpp:1990-2068 face weights; CG_benchmarks scenes).
 */
typedef struct uaamg_scene uaamg_scene;
int uaamg_scene_create(int config, int32_t n, uint64_t seed, uaamg_scene** out);
int uaamg_scene_create_custom(int32_t n, float pool, int32_t nspheres, const float* spheres, int32_t has_solid,
                              const float* solid_sphere, uint64_t velocity_seed, uaamg_scene** out);
/* Views into the scene's arrays (valid until destroy): phi, face weights, vx, vy, vz. */
int uaamg_scene_grids(const uaamg_scene* s, uaamg_grid* phi, uaamg_grid* weights, uaamg_grid* vx, uaamg_grid* vy,
                      uaamg_grid* vz, float* dx, int32_t* n);
void uaamg_scene_destroy(uaamg_scene* s);

/* fraction_inside of a segment / square (levelset_util.cpp:4-103), exposed for checks. */
float uaamg_fraction_inside2(float phi_left, float phi_right);
float uaamg_fraction_inside4(float phi_bl, float phi_br, float phi_tl, float phi_tr);

/* ------------------------------------------------------------- viscosity */
/*
 * Variational viscosity solve (simd_viscosity3d / L_with_level, SIMD_UAAMG/simd_viscosity3d.h:14-321).
 * Velocities are three face-centred channels: channel c of voxel v sits at world (v - e_c/2) dx
 * (packed_FloatGrid3, packed3grids.h:5-23). DOF-ordered vectors number channel 0's DOFs first, then
 * channel 1's, then channel 2's, each in leaf tree order and voxel offset order (set_dof, :504-537).
 * Grids: liquid SDF and viscosity on the voxel-centre lattice (index v at world v dx); solid SDF on the
 * vertex lattice (index i at world (i - 1/2) dx), as FLIP_vdb.cpp:2117-2131 sets them up.
 */
typedef struct uaamg_viscosity uaamg_viscosity;

typedef struct uaamg_viscosity_options {
    int32_t max_iterations;   /* m_max_iter, default 100 (simd_viscosity3d.cpp:3629) */
    int32_t device;           /* HIP device ordinal; -1 = host-only handle (inspection, no solves) */
    int32_t max_coarsest_dof; /* coarsen while ndof > this (4000, :3638) */
    int32_t flags;            /* 0 normally; bit 1 = no graph replay (A/B checks); bit 3 = keep level 0 untrimmed;
                                 bit 5 = small coarse levels (widest channel <= 16 leaves) as single-workgroup
                                 half-visits instead of the per-pass launches (the default); bit 8 = level-0
                                 volume stage (subcell SDF, subcell / voxel / face / edge volumes) on the host
                                 threads instead of the device (same result) */
    int32_t precision;        /* UAAMG_PRECISION_FP32 (default, the reference's arithmetic) or
                                 UAAMG_PRECISION_BF16_CYCLE: the W-cycle's level vectors below level 0 stored
                                 in BF16 and computed in FP32; level 0, the coarsest solve's arithmetic and
                                 the outer PCG stay FP32 with FP64 dot products. No reference counterpart. */
    void* stream;             /* hipStream_t to run on; NULL = a stream owned by the handle */
} uaamg_viscosity_options;

/* coefficient grids of a level (m_diagonal, m_invdiag, m_neg_x/y/z, m_cross_channel, m_SPAI0) */
#define UAAMG_VGRID_DIAG(c) (0 + (c))
#define UAAMG_VGRID_INVDIAG(c) (3 + (c))
#define UAAMG_VGRID_NEG(c, axis) (6 + 3 * (c) + (axis))
#define UAAMG_VGRID_CROSS(normal, j, k) (15 + 4 * (normal) + 2 * (j) + (k))
#define UAAMG_VGRID_SPAI0(c) (27 + (c))

/* light_weight_applier working modes (simd_viscosity3d.h:13) plus the two 6-colour sweeps */
#define UAAMG_VMODE_NORMAL 0
#define UAAMG_VMODE_RESIDUAL 1
#define UAAMG_VMODE_JACOBI 2
#define UAAMG_VMODE_SPAI0 3
#define UAAMG_VMODE_RED_GS 4
#define UAAMG_VMODE_BLACK_GS 5
#define UAAMG_VMODE_XYZ_RBGS 6
#define UAAMG_VMODE_ZYX_RBGS 7

int uaamg_viscosity_options_default(uaamg_viscosity_options* opts);

/*
 * Builds level 0 from the viscosity, liquid and solid grids, the coarse levels and the coarsest
 * explicit solver, and uploads them. Replaces L_with_level::create_viscosity_matrix
 * (simd_viscosity3d.h:95-98, .cpp:14-20) followed by the simd_viscosity3d constructor's hierarchy
 * loop (simd_viscosity3d.h:292-294, .cpp:3659-3690). dx is liquid_sdf->voxelSize()[0]. A problem
 * without degrees of freedom (no face-velocity DOF in the liquid) is rejected with UAAMG_EINVAL, as
 * uaamg_poisson_create rejects an empty liquid.
 */
int uaamg_viscosity_create(const uaamg_grid* viscosity, const uaamg_grid* liquid_sdf, const uaamg_grid* solid_sdf,
                           float dt, float rho, float dx, const uaamg_viscosity_options* opts,
                           uaamg_viscosity** out);
void uaamg_viscosity_destroy(uaamg_viscosity* h);
int uaamg_viscosity_set_options(uaamg_viscosity* h, const uaamg_viscosity_options* opts);

/* m_ndof of level 0 and m_matrix_levels.size() (simd_viscosity3d.h:225, 316). */
int64_t uaamg_viscosity_ndof(const uaamg_viscosity* h);
int uaamg_viscosity_nlevels(const uaamg_viscosity* h);

/* Per-level sizes: DOFs in total and per channel, leaves holding a DOF of any channel. */
int uaamg_viscosity_level_info(const uaamg_viscosity* h, int level, int64_t* ndof, int64_t* nchan3,
                               int64_t* nleaf_dof);

/*
 * DOF topology of a level (m_velocity_DOF / m_dof_manager[c], simd_viscosity3d.h:224-226): leaf
 * origins [nleaf_dof][3], per-channel masks [nleaf_dof][3][8] and per-channel first DOF index
 * [nleaf_dof][3] (-1 when the channel has no DOF in the leaf). Implements the data behind get_zero_vec
 * and the warm-start visitor of FLIP_vdb.cpp:3189-3197. Any output may be NULL.
 */
int uaamg_viscosity_level_topology(const uaamg_viscosity* h, int level, int32_t* origins, uint64_t* masks,
                                   int64_t* dofbase);

/*
 * Coefficient grid `grid` (UAAMG_VGRID_*) of a level as the level operator reads it, on the DOF leaves
 * ([nleaf_dof][512]), with kept[i] = 0 where the leaf is absent or trimmed to the default
 * (trim_default_nodes, :1700-1722).
 */
int uaamg_viscosity_level_coefficients(uaamg_viscosity* h, int level, int grid, float* bricks, uint8_t* kept);

/*
 * Explicit matrix of a level in CSR form with ascending columns (get_triplets / build_explicit_matrix,
 * :2598-2618). Call with ptr == NULL to get *nnz first.
 */
int uaamg_viscosity_level_matrix(const uaamg_viscosity* h, int level, int64_t* nnz, int64_t* ptr, int32_t* col,
                                 float* val);

/*
 * Right-hand side, DOF-ordered (L_with_level::build_rhs, simd_viscosity3d.h:128, .cpp:1942-2087).
 * velocity is a Vec3f grid (ncomp 3) of face-centred components; solid_velocity likewise or NULL (zero).
 */
int uaamg_viscosity_build_rhs(const uaamg_viscosity* h, const uaamg_grid* velocity, const uaamg_grid* solid_velocity,
                              float* rhs);

/*
 * simd_viscosity3d::pcg_solve(in_lhs, tolerance) (simd_viscosity3d.h:296, .cpp:3694-3740) on host
 * DOF-ordered vectors: x holds the initial guess on entry and the result on exit; *iterations receives
 * m_iteration. Returns 0 (the reference reports no failure status) or a negative error.
 */
/*
 * The same right-hand side with the liquid velocity as packed_FloatGrid3's three single-channel grids
 * v[0..2] (each read as its accessor would: the stored value where the channel has the leaf, else its
 * background). Replaces L_with_level::build_rhs (simd_viscosity3d.h:116, .cpp:1724-2087) as
 * simd_viscosity3d's constructor calls it. solid_velocity may be NULL (zero).
 */
int uaamg_viscosity_build_rhs_packed(const uaamg_viscosity* h, const uaamg_grid* v0, const uaamg_grid* v1,
                                     const uaamg_grid* v2, const uaamg_grid* solid_velocity, float* rhs);

/*
 * L_with_level::to_Eigenvector (simd_viscosity3d.cpp:2089-2104): vec[d] = channel c's stored value at
 * DOF d's face, for the DOFs of channel 0, then 1, then 2 (0 where the channel grid lacks the leaf).
 */
int uaamg_viscosity_grid_to_vector(const uaamg_viscosity* h, const uaamg_grid* v0, const uaamg_grid* v1,
                                   const uaamg_grid* v2, float* vec);

/*
 * L_with_level::write_to_FloatGrid3 (simd_viscosity3d.cpp:2123-2139) onto the level-0 DOF topology of
 * uaamg_viscosity_level_topology: bricks [nleaf_dof][3][512] receive vec at each channel's DOF faces
 * (active in that channel's grid) and 0 elsewhere. With all-zero vec this is get_zero_vec (:2156-2167).
 */
int uaamg_viscosity_vector_to_grid(const uaamg_viscosity* h, const float* vec, float* bricks);

/*
 * packed_FloatGrid3::from_vec3 (packed3grids.cpp:11-54) on flattened leaves: each channel c is the
 * Vec3f grid's topology (same origins, same active masks, same leaf order) with every value at
 * background[c], then component c at each active voxel. values: [3][nleaf][512]. The staggered
 * transform shift of the reference is the caller's (grid metadata).
 */
int uaamg_packed3_from_vec3(const uaamg_grid* vec3, float* values);

/*
 * packed_FloatGrid3::to_vec3 (packed3grids.cpp:57-83): the union of the three channels' leaves in tree
 * order, active = the union of their active masks, component c = channel c's value where channel c is
 * active and 0 elsewhere (background (0, 0, 0)). Call with origins = active = values = NULL to get
 * *nleaf; then pass arrays of *nleaf leaves (origins [n][3], active [n][8], values [n][512][3]).
 */
int uaamg_packed3_to_vec3(const uaamg_grid* v0, const uaamg_grid* v1, const uaamg_grid* v2, int64_t* nleaf,
                          int32_t* origins, uint64_t* active, float* values);

int uaamg_viscosity_solve_pcg(uaamg_viscosity* h, float tolerance, float* x, const float* b, int32_t* iterations);
/* Same with x and b in device memory. */
int uaamg_viscosity_solve_pcg_device(uaamg_viscosity* h, float tolerance, float* x_dev, const float* b_dev,
                                     int32_t* iterations);

/*
 * Merged viscosity batch: many simd_viscosity3d problems (uneven sizes, their own dt, rho, dx and
 * viscosity) merged into one leaf table per level, so every Gauss-Seidel colour pass, transfer and
 * PCG vector operation is one launch for the whole batch and one hipGraph per PCG iteration covers
 * every problem (no counterpart in the reference, which solves one problem per simd_viscosity3d;
 * each problem's result and m_iteration equal its pcg_solve alone, simd_viscosity3d.cpp:3694-3740).
 * FP32 form (opts->precision must be UAAMG_PRECISION_FP32); opts as uaamg_viscosity_create.
 */
typedef struct uaamg_viscosity_batch uaamg_viscosity_batch;
int uaamg_viscosity_batch_create(int32_t n, const uaamg_grid* const* viscosity, const uaamg_grid* const* liquid_sdf,
                                 const uaamg_grid* const* solid_sdf, const float* dt, const float* rho,
                                 const float* dx, const uaamg_viscosity_options* opts,
                                 uaamg_viscosity_batch** out);
void uaamg_viscosity_batch_destroy(uaamg_viscosity_batch* h);
/* problems, merged levels, merged level-0 leaves (with alignment padding), device bytes */
int uaamg_viscosity_batch_info(const uaamg_viscosity_batch* h, int32_t* nproblems, int32_t* nlevels,
                               int64_t* leaves0, int64_t* device_bytes);
/* m_ndof of problem i (-1 for a bad index) */
int64_t uaamg_viscosity_batch_ndof(const uaamg_viscosity_batch* h, int32_t i);
/* pcg_solve(tolerance) on every problem: x_dev[i] / b_dev[i] are DOF-ordered device vectors of
 * problem i (x: initial guess in, result out); iterations[i] its m_iteration (may be NULL). */
int uaamg_viscosity_batch_solve_device(uaamg_viscosity_batch* h, float tolerance, float* const* x_dev,
                                       const float* const* b_dev, int32_t* iterations);

/* Many independent viscosity solves at once, as uaamg_poisson_solve_batch_device. */
int uaamg_viscosity_solve_batch_device(uaamg_viscosity* const* handles, int32_t n, float tolerance,
                                       float* const* x_dev, const float* const* b_dev, int32_t* iterations,
                                       int32_t concurrency);

/*
 * Level operations on host DOF-ordered vectors: UAAMG_VMODE_* (L_apply, residual_apply, Jacobi_apply,
 * SPAI0_apply, XYZ_RBGS_apply, ZYX_RBGS_apply; simd_viscosity3d.cpp:2195-2302). RED_GS / BLACK_GS run
 * that colour pass on channels 0, 1, 2 in turn. b may be NULL for UAAMG_VMODE_NORMAL.
 */
int uaamg_viscosity_apply(uaamg_viscosity* h, int level, int mode, const float* x, const float* b, float* y);
/* restriction from fine_level to fine_level + 1 (:2353-2413). */
int uaamg_viscosity_restrict(uaamg_viscosity* h, int fine_level, const float* fine, float* coarse);
/* prolongation<true>: fine += coarse (:2659-2729). */
int uaamg_viscosity_prolong(uaamg_viscosity* h, int fine_level, const float* coarse, float* fine);
/* solve_exact on the coarsest level (:2645-2657). */
int uaamg_viscosity_coarsest_solve(uaamg_viscosity* h, const float* b, float* x);
/* One W-cycle from zero, z = M r (mucycle(z, r, 2, true, 0, 2), :3828-3884). */
int uaamg_viscosity_precondition(uaamg_viscosity* h, const float* r, float* z);

/* Measurement hooks for one colour / operator pass over channel c (no reference counterpart). */
int uaamg_viscosity_algorithmic_bytes(const uaamg_viscosity* h, int level, int mode, int channel, double* bytes);
/* Level-to-kernel mapping chosen when the handle was created: *small = 1 when the level's half-visits
 * run as one single-workgroup launch each, and the ms per half-visit pair of both mappings as timed
 * then (-1 when not timed). No reference counterpart. */
int uaamg_viscosity_kernel_map(const uaamg_viscosity* h, int level, int32_t* small, float* ms_small, float* ms_pass);
int uaamg_viscosity_time_apply(uaamg_viscosity* h, int level, int mode, int channel, int reps,
                               float* ms_per_launch);

/*
 * Synthetic viscosity scene (stand-in for FLIP's grids, FLIP_vdb.cpp:3170-3183): the liquid SDF of
 * scene `config`, the container walls (plus the solid sphere when the scene has one) as solid SDF,
 * constant viscosity mu, face velocities U[-1,1]; dt and rho as given.
 */
typedef struct uaamg_visc_scene uaamg_visc_scene;
int uaamg_visc_scene_create(int config, int32_t n, uint64_t seed, float mu, float dt, float rho,
                            uaamg_visc_scene** out);
int uaamg_visc_scene_grids(const uaamg_visc_scene* s, uaamg_grid* liquid_sdf, uaamg_grid* solid_sdf,
                           uaamg_grid* viscosity, uaamg_grid* velocity, float* dx, float* dt, float* rho);
void uaamg_visc_scene_destroy(uaamg_visc_scene* s);

/* volume_fraction of a cube from its 8 corner values (volume_fractions.cpp:100-126), for checks. */
float uaamg_volume_fraction8(const float* phi8);

/* Number of HIP devices visible (0 when none; never fails). */
int uaamg_device_count(void);

/* ================================================================ OpenVDB files
 * Reads an OpenVDB file (file format versions 222-224: OpenVDB 6-10, Houdini 18-20) without
 * OpenVDB, for the grids FLIP reads with openvdb::io::File (the liquid source and collision level
 * sets and their velocity grids, FLIP_vdb/main.cpp:108-131,161-185,192-213) and writes
 * (sdfvel<n>.vdb: liquid SDF and the three velocity channels; flip_particle<n>.vdb: the interior
 * SDF beside a point grid, FLIP_vdb/main.cpp:266-284). Float and Vec3f grids (full or half float;
 * uncompressed, zlib or Blosc with the LZ4 or zlib codec, active-mask compression) are decoded;
 * other grid types (point data, integer and mask grids) are listed with decoded = 0.
 */
typedef struct uaamg_vdb_file uaamg_vdb_file;
typedef struct uaamg_vdb_grid_desc {
    const char* name;        /* grid name (valid until uaamg_vdb_close) */
    const char* type;        /* tree type, e.g. "Tree_float_5_4_3" */
    const char* grid_class;  /* metadata "class": "level set", "fog volume", ... or "" */
    int32_t decoded;         /* 1: leaves available through uaamg_vdb_grid */
    int32_t ncomp;           /* 1 (float) or 3 (Vec3f); 0 for other types */
    int32_t half_float;      /* stored as half floats (read widened) */
    int64_t nleaf;           /* leaf nodes */
    int64_t ntiles;          /* root / internal tiles that are active or differ from the background */
    double voxel_size[3];    /* from the grid's transform */
    double translation[3];
    float background[3];
    int64_t meta_voxel_count; /* metadata "file_voxel_count" (-1 when absent) */
    int32_t meta_bbox[6];     /* metadata "file_bbox_min", "file_bbox_max" (index space) */
    int32_t has_meta_bbox;
} uaamg_vdb_grid_desc;
int uaamg_vdb_open(const char* path, uaamg_vdb_file** out);
void uaamg_vdb_close(uaamg_vdb_file* f);
int32_t uaamg_vdb_grid_count(const uaamg_vdb_file* f);
int uaamg_vdb_grid_info(const uaamg_vdb_file* f, int32_t index, uaamg_vdb_grid_desc* info);
/* The grid's leaves, in OpenVDB tree order, as a uaamg_grid whose arrays belong to the file handle
 * (valid until uaamg_vdb_close). Tiles are not expanded into leaves: see uaamg_vdb_grid_tiles. */
int uaamg_vdb_grid(const uaamg_vdb_file* f, int32_t index, uaamg_grid* out);
/* The tiles: origins [ntiles][3], edge in voxels [ntiles] (8, 128 or 4096), values [ntiles][3],
 * active [ntiles]; any array may be NULL. */
int uaamg_vdb_grid_tiles(const uaamg_vdb_file* f, int32_t index, int32_t* origins, int32_t* dims, float* values,
                         int32_t* active);
/* Turns the grid's tiles of edge <= max_dim (8, 128 or 4096) into leaves holding the tile value
 * (all voxels active when the tile is), in tree order, so that the solvers' lookups outside the
 * leaves see what an OpenVDB accessor returns there (a level set's interior tiles: -background);
 * *added receives the number of leaves added. UAAMG_ENOMEM (nothing changed) when the grid would
 * exceed max_leaves. Earlier uaamg_vdb_grid views of the grid are invalidated. */
int uaamg_vdb_expand_tiles(uaamg_vdb_file* f, int32_t index, int32_t max_dim, int64_t max_leaves, int64_t* added);
/* The Blosc 1 decoder the reader uses (LZ4 or zlib codec, byte shuffle): decompresses the Blosc
 * buffer src[srclen] into exactly dstlen bytes. 0 or UAAMG_EINVAL (malformed or unsupported). */
int uaamg_blosc_decompress(const void* src, int64_t srclen, void* dst, int64_t dstlen);

/* ================================================================ solve server
 * A request loop over many resident problems on one GPU (no reference counterpart: FLIP_vdb
 * builds a solver per time step and calls it synchronously, FLIP_vdb.cpp:3126-3208). Problems
 * are registered once from the same inputs as uaamg_poisson_create / uaamg_viscosity_create;
 * the server keeps their host hierarchies and makes a problem resident in device memory
 * (hierarchy, PCG vectors, warm-start solution) when one of its requests is admitted. Each of
 * `slots` worker threads takes the oldest queued request whose problem is idle and whose
 * footprint (uaamg_server_problem_info's device_bytes, summed over the levels from their
 * active-leaf counts) fits the memory budget, evicting idle resident problems least recently
 * used first; an evicted problem's solution is kept on the host. Problems own their streams,
 * so admitted solves overlap on the device; the requests of one problem run in submission order.
 */
typedef struct uaamg_server uaamg_server;

typedef struct uaamg_server_options {
    int32_t device;              /* HIP device ordinal */
    int32_t slots;               /* worker threads = solves in flight at most, 1..16 (default 4) */
    int64_t memory_budget_bytes; /* device bytes for resident problems; 0 = 90% of free memory */
} uaamg_server_options;

typedef struct uaamg_server_counters {
    int64_t completed, failed, evictions, acquisitions, resident_bytes, peak_inflight, queued;
} uaamg_server_counters;

int uaamg_server_options_default(uaamg_server_options* opts);
int uaamg_server_create(const uaamg_server_options* opts, uaamg_server** out);
/* Finishes every queued request, then frees every problem. */
void uaamg_server_destroy(uaamg_server* s);

/* Registers a problem (the hierarchy is built on the host here); *problem receives its id. The
 * options' stream and device fields are replaced by the problem's own stream on the server's
 * device. */
int uaamg_server_add_poisson(uaamg_server* s, const uaamg_grid* liquid_phi, const uaamg_grid* face_weights, float dt,
                             float dx, const uaamg_options* opts, int64_t* problem);
int uaamg_server_add_viscosity(uaamg_server* s, const uaamg_grid* viscosity, const uaamg_grid* liquid_sdf,
                               const uaamg_grid* solid_sdf, float dt, float rho, float dx,
                               const uaamg_viscosity_options* opts, int64_t* problem);
int uaamg_server_problem_info(uaamg_server* s, int64_t problem, int64_t* ndof, int64_t* device_bytes,
                              int32_t* resident);

/* Queues one solve to `tolerance` (relative, as the problem's solver defines it). b_dev: the
 * right-hand side (device, DOF order); x_out_dev: where the solution is copied (device, NULL
 * for none); x0_dev: the initial guess (device; NULL = the problem's previous solution, zero
 * before its first solve). The caller keeps the three buffers alive until uaamg_server_wait
 * returns for the ticket. UAAMG_ENOMEM when the problem can never fit the budget. */
int uaamg_server_submit(uaamg_server* s, int64_t problem, const float* b_dev, float* x_out_dev, const float* x0_dev,
                        float tolerance, int64_t* ticket);
/* Blocks until the ticket's solve completes. *status receives the solver's status (UAAMG_SUCCESS,
 * UAAMG_FAILED or a negative error), the times are in ms from submission (queueing, and total).
 * After a negative *status, uaamg_last_error() on the same thread names the failed step and its
 * HIP error. */
int uaamg_server_wait(uaamg_server* s, int64_t ticket, int32_t* status, int32_t* iterations, double* queue_ms,
                      double* total_ms);
/* uaamg_server_wait plus the pure-multigrid fallback of a Poisson problem registered with flags bit 7:
 * when its PCG returned FAILED (*status = UAAMG_FAILED, *iterations its PCG iterations), the server
 * zeroed the solution and ran solvePureMultigrid (FLIP_vdb.cpp:3150-3155); *mg_status and
 * *mg_iterations receive that run's status and iterations, -1 when it did not run. */
int uaamg_server_wait_fallback(uaamg_server* s, int64_t ticket, int32_t* status, int32_t* iterations,
                               int32_t* mg_status, int32_t* mg_iterations, double* queue_ms, double* total_ms);
int uaamg_server_stats(uaamg_server* s, uaamg_server_counters* out);

#ifdef __cplusplus
}
#endif

#endif /* UAAMG_CAPI_H */
