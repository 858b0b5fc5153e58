/*
 * naja_amd — C ABI of the MI355X (gfx950) coordinate hit-and-run sampler.
 *
 * Every entry point takes plain pointers and sizes. Matrix and vector
 * arguments are DEVICE pointers (hipMalloc'd or torch CUDA tensors) unless a
 * comment says "host". `stream` is a hipStream_t passed as void* (NULL = the
 * null stream). Calls are asynchronous on that stream; they return after the
 * launches are queued. Return value: 0 on success or a negative NAJA_ERR_*
 * code; naja_last_error() then holds a thread-local message. No C++ exception
 * crosses this boundary.
 *
 * The reference boundary these replace is the C++ family in naja::gpu,
 * src/gpu/gpusamplers.h:30-86, which takes DMatrix<double>/DVector<double>
 * device containers and throws std::invalid_argument / std::runtime_error.
 */
#ifndef NAJA_AMD_H
#define NAJA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NAJA_OK 0
#define NAJA_ERR_INVALID_ARGUMENT (-1) /* reference: std::invalid_argument (backmap.cu:27-35) */
#define NAJA_ERR_RUNTIME (-2)          /* reference: std::runtime_error (chr.cu:549) */
#define NAJA_ERR_HIP (-3)              /* HIP API failure (the reference only prints: helper.h:12-27) */
#define NAJA_ERR_UNSUPPORTED (-4)      /* shape outside the compiled kernel variants */

/* Sample layout selectors for naja_diag. */
#define NAJA_LAYOUT_INTERLEAVED 0 /* row s*nchains + c is sample s of chain c (samples.npy as written, npy.h:96-141) */
#define NAJA_LAYOUT_BLOCKS 1      /* rows [c*spc, (c+1)*spc) are chain c (naja validate's grouping, mcmc.cpp:103-110) */

/* Thread-local message describing the last failed call on this thread. */
const char* naja_last_error(void);

/* Library ABI version (major*10000 + minor*100 + patch). */
int naja_version(void);

/*
 * Sampler parameters. Field meanings follow the arguments of
 * naja::gpu::CoordinateHitAndRun (gpusamplers.h:30-48) and the
 * `naja sample run` flags (cmd_run.cpp:76-103).
 */
typedef struct naja_chr_params {
    int nspc;             /* samples per chain (reference `nspc`, --n-samples) */
    int thinning;         /* steps between saved samples (--thinning; > 0) */
    int nchains;          /* chains per polytope / per variant (--n-chains) */
    int tpb;              /* reference --tpb: workgroup size request; <= 0 lets the library choose */
    uint64_t seed;        /* Philox key (the reference's int seed, widened) */
    double pair_prob;     /* --pair-prob */
    int resync_interval;  /* --resync-interval: FP64 s = b - A y every R steps; 0 = automatic
                             max(1024, 16 d) window; < 0 = never after t = 0 */
    double ksparse_prob;  /* --ksparse-prob */
    int ksparse_k;        /* --ksparse-k (clamped to [1, min(32, d)] as chr.cu:300-304) */
    int pair_mode;        /* 1 = (e_i - e_j)/sqrt2, 2 = schedule pairs (plan.h:13-19) */
    const int* pair_i;    /* device, n_pairs (pair mode 2; schedule_io.cpp:12-75 columns i, j, c, s) */
    const int* pair_j;
    const float* pair_c;
    const float* pair_s;
    int n_pairs;
    int chains_per_group; /* G chains share one move draw (1 = independent moves, as naja) */
    int32_t* move_log;    /* device, optional: (type, i, j) of chain 0's moves for t < nlog */
    int64_t nlog;
} naja_chr_params;

/*
 * naja::gpu::CoordinateHitAndRun (gpusamplers.h:30-48, chr.cu:13-70).
 *   A       : m x d, column-major (lda = m), FP32
 *   b       : m, FP64 (used by the FP64 slack computation s = b - A y)
 *   y0      : d x nchains column-major start points (the reference's X_d)
 *   samples : d x (nchains*nspc) column-major; column s*nchains + c holds
 *             sample s of chain c (chr.cu:284, chr.cu:460). Read as a C-order
 *             (nchains*nspc, d) array it is samples.npy (npy.h:96-141).
 *   y_final : optional d x nchains final states (the reference's
 *             persist_state write-back, chr.cu:485-489).
 */
int naja_chr_sample(const float* A, const double* b, int m, int d,
                    const float* y0, const naja_chr_params* p,
                    float* samples, float* y_final, void* stream);

/*
 * Batched conditioned variants in one persistent launch: variant v samples
 * {y : [A; E_v] y <= [b; f_v]} with E_v its own k_v extra rows (the
 * reference stacks one extra block per model: extra_constraints.cpp:17-58,
 * and runs models one job at a time: job_bulk.cpp:115-189).
 *   extra_rows : HOST, n_variants counts k_v >= 0
 *   A_extra    : concatenated column-major k_v x d blocks, variant order
 *   b_extra    : concatenated k_v right-hand sides (FP64)
 *   y0         : d x (n_variants*nchains), variant-major
 *   samples    : n_variants consecutive d x (nchains*nspc) blocks
 * Variant v uses Philox key (seed, variant v).
 */
int naja_chr_sample_batched(const float* A, const double* b, int m0, int d,
                            int n_variants, const int* extra_rows,
                            const float* A_extra, const double* b_extra,
                            const float* y0, const naja_chr_params* p,
                            float* samples, float* y_final, void* stream);

/*
 * Batched variants with an optional rounding transform per variant: variant
 * v samples z on {z : [A; E_v] (c_v + P_v z) <= [b; f_v]} and returns
 * y = c_v + P_v z (a fixed linear map keeps the uniform target uniform).
 *   P : n_variants x d x d row-major (NULL for none), c : n_variants x d
 *   y0: z-space starts, d x (n_variants*nchains) (z = 0 is y = c_v)
 *   y_final: final z-space states
 * No counterpart in the reference (naja reads an already rounded polytope,
 * runtime_config.h:64-78); this is how the Dikin and learned preconditioners
 * are applied per conditioned variant.
 */
int naja_chr_sample_batched_rounded(const float* A, const double* b, int m0, int d,
                                    int n_variants, const int* extra_rows,
                                    const float* A_extra, const double* b_extra,
                                    const float* P, const float* c,
                                    const float* y0, const naja_chr_params* p,
                                    float* samples, float* y_final, void* stream);

/*
 * Sampling plans: pack (and transform) the polytope blocks once, then launch
 * any number of times without allocating, so a launch can be captured in a
 * hipGraph. A plan keeps pointers to A, b, A_extra, b_extra, P and c (the
 * caller keeps them alive) and allocates, packs and frees its own buffers on
 * the creation stream. Launches and autotuning may use another stream: that
 * stream first waits (once per repack) for the packing, and the creation
 * stream waits for such launches before it repacks or frees. A captured
 * launch records neither wait: capture a plan only after its creation (or
 * autotuning) stream has finished packing (e.g. after a synchronisation),
 * and order graph replays before naja_chr_plan_destroy or a later autotune
 * yourself. naja_chr_plan_autotune times every compiled launch shape
 * on about 1024 steps of the real run, keeps the fastest, and remembers it
 * for later plans of the same size (results do not depend on the shape).
 * shape_out receives {wavefronts per workgroup, chains per group, rows per
 * lane}; in the paired-row layout rows per lane count packed rows. Creation
 * checks every block (A and each variant's extra rows) for a second half
 * that negates the first ([T; -T] bounds, [a; -a] slabs) and packs such
 * blocks one row per pair, two slacks per packed row; the check reads one
 * flag per block back, so creation synchronises its stream once. Samples
 * are those of the row-by-row algorithm in either layout. The one-shot
 * sampling calls create a plan internally and synchronise likewise.
 * The reference builds its launch on every call (chr.cu:13-70) and has no
 * autotuner (launchChrKernel takes --tpb as given, chr.cu:499-553).
 */
typedef struct naja_chr_plan naja_chr_plan;
int naja_chr_plan_create(const float* A, const double* b, int m0, int d,
                         int n_variants, const int* extra_rows,
                         const float* A_extra, const double* b_extra,
                         const float* P, const float* c,
                         const naja_chr_params* p, void* stream, naja_chr_plan** out);
/*
 * Replaces the rounding of a plan created with one (P, c as at creation:
 * n_variants x d x d row-major and n_variants x d, device memory the caller
 * keeps alive): the packed rows stay and only A' = A P, b' = b - A c are
 * formed again, on the creation stream, with no allocation and no
 * synchronisation; later launches sample the new rounding. A, b and the
 * extra rows must be unchanged. For pipelines that recompute the rounding of
 * the same conditioned variants at every step (the reference rounds a model
 * once, offline: src/rounding/plan.cpp:12-62).
 */
int naja_chr_plan_set_rounding(naja_chr_plan* plan, const float* P, const float* c);
int naja_chr_plan_autotune(naja_chr_plan* plan, const float* y0, void* stream, int* shape_out);
int naja_chr_plan_shape(const naja_chr_plan* plan, int* shape_out);
/*
 * The row layout the plan's launches run: layout_out receives {1 if the
 * paired-row layout, chains per workgroup (4 when the packed rows are read
 * from LDS, else 1), 1 if the packs hold scaled pairs (every block paired:
 * each pair's rows and right-hand sides divided by a power of two of at least
 * twice its width, reciprocals multiplied by it)}. Samples do not depend on
 * it. No counterpart in the reference (its launch has one row layout,
 * chr.cu:13-70); for callers and tests that check which kernel a plan runs.
 */
int naja_chr_plan_layout(const naja_chr_plan* plan, int* layout_out);
int naja_chr_plan_launch(naja_chr_plan* plan, const float* y0, float* samples, float* y_final, void* stream);
int naja_chr_plan_destroy(naja_chr_plan* plan);

/*
 * Streamed sampling (CoordinateHitAndRunStreamed, chr.cu:101-230): the
 * chains run in chunks of chunk_nspc samples; each chunk starts from Y
 * (nchains x d device FP32, updated in place: the reference's persist_state)
 * at the global step chunk_start * thinning, so the Philox stream continues
 * and the chunks reproduce one long run bit for bit when chunk_nspc *
 * thinning is a multiple of the effective resync interval. With T (n x d)
 * and shift (n) each chunk is backmapped on the device. Chunk k is copied to
 * pinned host memory on a second stream while chunk k + 1 computes, then
 * handed to sink(rows, n_rows, n_cols, first_row, user) with rows in
 * samples.npy order (row s * nchains + c); a non-zero sink return aborts.
 */
typedef int (*naja_sample_sink)(const float* rows, int64_t n_rows, int n_cols, int64_t first_row, void* user);
int naja_chr_sample_streamed(const float* A, const double* b, int m, int d, float* Y, const naja_chr_params* params,
                             int chunk_nspc, const float* T, const float* shift, int n, naja_sample_sink sink,
                             void* user, void* stream);

/*
 * backmap_into (backmap.cu:21-48): V = T Y + shift, FP32 MFMA.
 *   T     : n x d row-major (reaction rows, as the CSV stores T)
 *   shift : n
 *   Y     : ns x d row-major (sample rows)
 *   V     : ns x n row-major (backmapped sample rows)
 */
int naja_backmap(const float* T, const float* shift, const float* Y, float* V,
                 int n, int d, int64_t ns, void* stream);

/*
 * compute_ess + compute_split_rhat (mcmc.cpp:91-194) on the GPU, for chains
 * of any length (up to 2^27 samples): transforms of up to 2^15 points run in
 * LDS, longer ones in global memory.
 *   samples   : n_total x dim row-major (samples.npy), n_total % nchains == 0
 *   layout    : NAJA_LAYOUT_INTERLEAVED or NAJA_LAYOUT_BLOCKS
 *   ess       : dim, per-dim minimum over chains
 *   rhat      : dim, classic split-R-hat
 *   ess_chain : optional dim x nchains per-chain ESS
 */
int naja_diag(const float* samples, int dim, int n_total, int nchains, int layout,
              float* ess, float* rhat, float* ess_chain, void* stream);

/*
 * naja_diag without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments and
 * results as naja_diag. Its scratch (the transposed series, the half-chain
 * moments and, when ess_chain is NULL, the per-chain ESS: about
 * 4 * dim * nchains * (spc rounded up to 64) bytes plus 32 * dim * nchains)
 * lives in a static device buffer of 64 MiB per GPU, held until the process
 * ends; larger calls fail with NAJA_ERR_UNSUPPORTED. Every call on one GPU
 * shares that buffer, so calls that overlap on one GPU (different streams or
 * host threads) race. The first call on a device sets it up with a
 * synchronous 7 KB upload of the FFT twiddle tables, so it must not be made
 * under stream capture; later calls only launch kernels on `stream`.
 */
int naja_diag_noalloc(const float* samples, int dim, int n_total, int nchains, int layout,
                      float* ess, float* rhat, float* ess_chain, void* stream);

/*
 * naja_diag over n_variants sample sets stored back to back (variant v at
 * samples + v * n_total * dim), in one set of launches. ess, rhat:
 * n_variants x dim; ess_min, rhat_max (nullable): per variant, the minimum
 * over dims of ess and the maximum of rhat. No reference counterpart (the
 * reference validates one sample file at a time, validate_cli.cpp:49-296).
 */
int naja_diag_batched(const float* samples, int n_variants, int dim, int n_total, int nchains, int layout,
                      float* ess, float* rhat, float* ess_min, float* rhat_max, void* stream);

/*
 * Measurement hook (no reference counterpart; for benchmarks): while enabled,
 * every CHR step-kernel launch records a pair of HIP events on its stream.
 * naja_chr_kernel_times waits for and returns up to max recorded launches
 * (work items, i.e. chain groups, and milliseconds), oldest first, writes how
 * many it returned to *count, and clears the record. Off by default.
 */
int naja_chr_timing(int enable);
int naja_chr_kernel_times(int* items, float* ms, int max, int* count);

/* ---------------------------------------------------------------------------
 * Affine-hull reduction (host, FP64, row-major host memory), replacing the
 * Eigen::FullPivLU block of run_sampling_job (job_sampling.cu:196-235).
 * ------------------------------------------------------------------------- */

/*
 * Kernel basis of Aeq (r x n) by full-pivot LU with relative pivot threshold
 * `threshold` (a pivot counts when |u_kk| > threshold * max |pivot|; the
 * reference uses 1e-10, job_sampling.cu:211). basis: caller buffer of n*n
 * doubles; receives n x dim row-major (leading dimension dim), one column per
 * non-pivot column f of the permuted matrix, with entry 1 at that original
 * column and 0 at the other free columns. r = 0 gives the identity.
 */
int naja_kernel_basis(const double* Aeq, int r, int n, double threshold, double* basis, int* dim_out);

/*
 * The equality set of {x : A x <= b} at x0 is the rows with b_i - a_i.x0 <= tol
 * (job_sampling.cu:199-204); basis (n*n buffer) receives its kernel basis as in
 * naja_kernel_basis, dim_out its dimension, n_tight_out (nullable) the row
 * count. The caller reduces when 0 < dim < n: A' = A basis, b' = b - A x0,
 * z0 = 0, x = x0 + basis z (job_sampling.cu:213-223).
 */
int naja_affine_hull(const double* A, const double* b, const double* x0, int m, int n, double tol,
                     double threshold, double* basis, int* dim_out, int* n_tight_out);

/* ---------------------------------------------------------------------------
 * Producers of conditioned variants (host, FP64, row-major host memory).
 * ------------------------------------------------------------------------- */

/*
 * Start point of a polytope: the centre of the largest axis-aligned cube
 * inside {x : A x <= b}, i.e. the LP  max r  s.t.  A x + ||a_i||_1 r <= b,
 * r >= 0 (axis_aligned_cube_center_lp_gurobi, feasible_start_lp.cpp:33-87,
 * which calls Gurobi). Own interior-point solver; x_hint (nullable) seeds it.
 * r_out is the exact radius min_i (b_i - a_i.x) / ||a_i||_1 at the returned x.
 * When the optimal centre is not unique, x is an interior point of the
 * optimal face (the reference returns Gurobi's choice). Errors:
 * NAJA_ERR_INVALID_ARGUMENT for an empty or non-finite input,
 * NAJA_ERR_RUNTIME for an infeasible or unbounded LP (the reference throws
 * runtime_error "LP status ..."). nthreads <= 0: all host threads.
 */
int naja_cube_center(const double* A, const double* b, int m, int n, const double* x_hint, double* x_out,
                     double* r_out, int nthreads);

/*
 * build_extra_from_bounds (cmd_prepare.cpp:111-156): for each flux i of the
 * rounded model x = T y + shift (T n x d), a tightened upper bound
 * (ub_new < ub_base - eps) adds the row T[i,:] with rhs ub_new - shift, a
 * tightened lower bound (lb_new > lb_base + eps) the row -T[i,:] with rhs
 * -lb_new + shift, in flux order (the reference uses eps = 1e-9). Both bound
 * sets are checked as validate_bounds does (cmd_prepare.cpp:80-93).
 * extra_A: caller buffer of 2 n d doubles (k x d row-major out); extra_b: 2 n.
 */
int naja_extra_from_bounds(const double* T, const double* shift, int n, int d, const double* lb_base,
                           const double* ub_base, const double* lb_new, const double* ub_new, double eps,
                           double* extra_A, double* extra_b, int* k_out);

/* ---------------------------------------------------------------------------
 * Rounding preconditioners (host, FP64, row-major host memory). Every P
 * returned here is the sampler's rounding transform: y = x_c + P z.
 * ------------------------------------------------------------------------- */

/*
 * Symmetric eigendecomposition (Householder tridiagonalisation + implicit QL),
 * standing in for Eigen::SelfAdjointEigenSolver (dikin_precondition.cpp:103).
 * a: n x n symmetric; evals ascending; evecs n x n with eigenvector j in column j.
 */
int naja_symmetric_eigen(const double* a, int n, double* evals, double* evecs);

/*
 * dikin_precondition (dikin_precondition.cpp:36-183, defaults in
 * dikin_precondition.h:35-42: k_directions 64, tight_factor 100,
 * target_chord 0 = median coordinate chord, alpha_max 1000).
 *   A m x d, b m, x_c d (NULL = origin).
 *   Outputs (each nullable): P (d x d, the reference's DikinResult::P),
 *   P_inv (d x d, y = x_c + P_inv z), A_new = A P_inv (m x d),
 *   b_new = b - A x_c (m), n_corrected.
 * Empty polytope (m or d = 0): NAJA_ERR_INVALID_ARGUMENT (the reference throws
 * std::invalid_argument).
 */
int naja_precond_dikin(const double* A, const double* b, const double* x_c, int m, int d,
                       int k_directions, double tight_factor, double target_chord, double alpha_max,
                       double* P, double* P_inv, double* A_new, double* b_new, int* n_corrected);

/*
 * dikin_precondition for every conditioned variant [A0; E_v] y <= [b0; f_v]
 * (the extra-constraint stacking of extra_constraints.cpp:17-58), variants
 * spread over host threads (nthreads <= 0: all cores).
 *   extra_rows[v] = k_v; extra_A: sum k_v x d rows, variant after variant;
 *   extra_b: sum k_v; x_c: n_variants x d (NULL = origin).
 *   P_inv: n_variants x d x d; n_corrected: n_variants (nullable).
 */
int naja_precond_dikin_batched(const double* A0, const double* b0, int m0, int d, int n_variants,
                               const int* extra_rows, const double* extra_A, const double* extra_b,
                               const double* x_c, int k_directions, double tight_factor,
                               double target_chord, double alpha_max, double* P_inv,
                               int* n_corrected, int nthreads);

/*
 * Dikin-ellipsoid whitening per variant: H = sum over all rows of
 * a a^T / s^2 at x_c, H = L L^T, P = L^{-T} (so P P^T = H^{-1}). No reference
 * counterpart: the frame the learned preconditioner's pilot chains run in.
 * x_c must be interior (every slack > 0), else NAJA_ERR_INVALID_ARGUMENT.
 */
int naja_precond_ellipsoid_batched(const double* A0, const double* b0, int m0, int d, int n_variants,
                                   const int* extra_rows, const double* extra_A, const double* extra_b,
                                   const double* x_c, double* P, int nthreads);

/*
 * naja_precond_ellipsoid_batched on the GPU, FP64 inside, for d x d and the
 * variant's rows in LDS ((d^2 + d + rows) x 8 B <= 150 KB, else
 * NAJA_ERR_UNSUPPORTED). Device memory, the batched-sampler layout:
 * A0 column-major (m0 x d), b0 FP64, extra_A the variants' column-major
 * k_v x d blocks back to back, extra_b FP64; extra_rows on the host; x_c
 * (nullable) n_variants x d; P out: n_variants x d x d FP32 row-major.
 * Synchronises the stream (per-variant status is checked).
 */
int naja_precond_ellipsoid_device(const float* A0, const double* b0, int m0, int d, int n_variants,
                                  const int* extra_rows, const float* extra_A, const double* extra_b,
                                  const float* x_c, float* P, void* stream);

/*
 * Dikin-ellipsoid rounding of one polytope of any size on the GPU (FP64):
 * the path beyond naja_precond_ellipsoid_device's LDS budget (genome scale
 * d = 512, large scale d = 2400 with 21200 rows). H = A^T diag(1/s^2) A with
 * s = b - A x_c (rocBLAS dsyrk on the FP64 MFMA units), a blocked Cholesky
 * H = L L^T (64 x 64 diagonal blocks in LDS, panels and trailing updates by
 * rocBLAS dtrsm / dsyrk), P = L^{-T}, so P P^T = H^{-1}. No reference
 * counterpart (the reference's Dikin rounding, dikin_precondition.cpp:36-183,
 * corrects only the stiffest directions).
 *   A m x d column-major FP32, b m FP64, x_c d FP32 (NULL = origin), all
 *   device; P out: d x d row-major FP32, upper triangular.
 * Scratch 8 (m d + 2 d^2) bytes. Synchronises the stream.
 * NAJA_ERR_INVALID_ARGUMENT for a centre that is not interior.
 */
int naja_precond_ellipsoid_large(const float* A, const double* b, int m, int d, const float* x_c, float* P,
                                 void* stream);

/*
 * Analytic centre of {y : A y <= b} (the minimiser of -sum log(b - A y)) on
 * the GPU by damped Newton steps from the interior point x0 (NULL = origin):
 * gradient and Hessian A^T diag(1/s^2) A in FP64 (rocBLAS dgemv / dsyrk),
 * the blocked Cholesky of naja_precond_ellipsoid_large, step 1 / (1 + lambda)
 * while the Newton decrement lambda >= 1/4 (the iterate stays interior), full
 * steps after; stops when lambda^2 < tol. The centre to round a polytope at
 * when its start sits close to some facets (the large-scale polytopes' start
 * is 1e-3 from its closest facet). A m x d column-major FP32, b m FP64, x0
 * and x_out d FP32, all device; iters_out (host, nullable) the steps taken.
 * NAJA_ERR_INVALID_ARGUMENT for a start that is not interior,
 * NAJA_ERR_RUNTIME when tol is not met within max_iter. Synchronises the
 * stream. No reference counterpart (naja starts at the rounding's
 * start.csv, runtime_config.h:64-78).
 */
int naja_analytic_center(const float* A, const double* b, int m, int d, const float* x0, int max_iter, double tol,
                         float* x_out, int* iters_out, void* stream);

/*
 * Facet-aligned rounding of one polytope with at least d independent facets,
 * at any size on the GPU (FP64; the large-scale companion of
 * naja_precond_facet_device): the dual rows a_r / s_r (s = b - A x_c) in
 * order of tightness ||a_r|| / s_r, each accepted while its residual against
 * the accepted ones keeps at least theta of its norm (so the other side of
 * an accepted slab, parallel to it, is skipped), until d are accepted;
 * P = N^-1 for the accepted rows N, so every accepted facet is a coordinate
 * slab of z. Blocks of 128 candidates projected by rocBLAS dgemm (two
 * Gram-Schmidt passes), selected in LDS, P = Q R^-1 by dtrsm. No reference
 * counterpart.
 *   A m x d column-major FP32, b m FP64, x_c d FP32 (NULL = origin), all
 *   device; P out: d x d row-major FP32.
 * NAJA_ERR_INVALID_ARGUMENT for a centre that is not interior,
 * NAJA_ERR_RUNTIME when fewer than d facets are independent.
 * Synchronises the stream.
 */
int naja_precond_facet_large(const float* A, const double* b, int m, int d, const float* x_c, float theta, float* P,
                             void* stream);

/* How naja_precond_facet_device completes the directions the facets leave free. */
#define NAJA_FACET_COMPLETE_ROUNDING 0    /* the dual rows of the base rounding P0 (rows of P0^-1) */
#define NAJA_FACET_COMPLETE_LOWER 1       /* rows of L^-1, L the lower Cholesky factor of P0 P0^T */
#define NAJA_FACET_COMPLETE_BASE_FACETS 2 /* the base polytope's dual rows, orthonormalised; P0 unused */

/*
 * Facet-aligned rounding per variant on the GPU: y = x_c + P z with the
 * coordinate axes of z along the edges of a parallelotope of binding facets,
 * where coordinate hit-and-run draws exactly (an axis-aligned box) instead of
 * merely whitened. Candidates are the dual rows a_r / s_r (s_r the slack at
 * x_c) of the variant's extra rows (with include_base, of its base rows too),
 * chosen greedily by modified Gram-Schmidt residual norm while the best one
 * keeps at least theta[v] of its own norm. The remaining directions are
 * completed by `completion`:
 *   NAJA_FACET_COMPLETE_ROUNDING: the dual rows of the base rounding P0 (the
 *     rows of P0^-1), greedily by residual;
 *   NAJA_FACET_COMPLETE_LOWER: the rows of L^-1 for the lower Cholesky factor
 *     L of P0 P0^T (the same ellipsoid with its axes in triangular order);
 *   NAJA_FACET_COMPLETE_BASE_FACETS: the base polytope's own dual rows, their
 *     residuals against everything chosen, the largest first, each made
 *     orthonormal (FP64 Gram-Schmidt) and kept as a row of N: the free axes
 *     lie in the subspace the chosen facets leave free, as close to the base
 *     facets' normals as an orthonormal basis of it gets. P0 is not read (may
 *     be NULL); include_base must be 0.
 * exchange_tol (0 = off, else >= 1): when the chosen facets fill the basis,
 * a facet candidate replaces a basis row while its coefficient in the basis
 * exceeds exchange_tol in magnitude (maximum-volume exchange: each step
 * shrinks the parallelotope {|N y| <= 1}). P = N^-1 for the d chosen rows N
 * (FP64 Gauss-Jordan). theta[v] > 1 selects no facet (with completion
 * NAJA_FACET_COMPLETE_ROUNDING, P0 up to the order of its columns). No
 * reference counterpart (naja samples an already rounded polytope,
 * runtime_config.h:64-78).
 * Device memory, the batched-sampler layout as naja_precond_ellipsoid_device:
 * A0 column-major (m0 x d), b0 FP64, extra_A the variants' column-major
 * k_v x d blocks, extra_b FP64; extra_rows on the host; x_c (nullable)
 * n_variants x d; P0 and P n_variants x d x d FP32 row-major; theta
 * n_variants floats. d <= 128 and at most 512 candidates (facets plus d, or
 * plus m0 with NAJA_FACET_COMPLETE_BASE_FACETS), else NAJA_ERR_UNSUPPORTED.
 * Synchronises the stream (per-variant status: NAJA_ERR_INVALID_ARGUMENT for
 * a centre that is not interior, NAJA_ERR_RUNTIME for a singular P0 or no
 * independent completion).
 */
int naja_precond_facet_device(const float* A0, const double* b0, int m0, int d, int n_variants,
                              const int* extra_rows, const float* extra_A, const double* extra_b,
                              const float* x_c, const float* P0, const float* theta, int include_base,
                              int completion, float exchange_tol, float* P, void* stream);

/*
 * Conditioned-variant run as one persistent hipGraph (no reference
 * counterpart: naja samples one model per job, job_sampling.cu:315-379, and
 * rounds it offline). A step is: the facet-aligned rounding of every variant
 * (naja_precond_facet_device with NAJA_FACET_COMPLETE_BASE_FACETS, theta,
 * exchange_tol, centre at the origin), the variants' rows packed again from
 * A0, b0, extra_A, extra_b and the rounding set on them
 * (naja_chr_plan_set_rounding), the batched CHR launch of every variant from
 * z = y0 with its backmap y = c + P z into `samples`, and the batched ESS and
 * split-R-hat (naja_diag_batched, interleaved layout) into ess, rhat (per
 * variant and dim), ess_min and rhat_max (per variant). Creation allocates
 * everything the step uses (tables, status words, the plan's packs, the
 * diagnostics workspace and twiddle tables), runs one step eagerly, then
 * captures the step on `stream` and instantiates the graph; so creation
 * synchronises `stream`, and the first step's results are in the outputs
 * when it returns. naja_cond_run_step replays the graph on `stream` (any
 * stream), copies the roundings' status back to pinned host memory and
 * synchronises: a variant whose centre is not interior gives
 * NAJA_ERR_INVALID_ARGUMENT, one without an independent completion
 * NAJA_ERR_RUNTIME. Replays are bit-identical to the eager step (same
 * kernels, same Philox stream). With timed != 0 the graph carries external
 * event nodes around the CHR launch and naja_cond_run_sampling_ms returns its
 * duration in the last step. Device memory, the layouts of
 * naja_chr_plan_create (A0, b0, extra_A, extra_b), y0 n_variants x nchains x
 * d, c n_variants x d, P (written) n_variants x d x d, samples n_variants x
 * (nchains nspc) x d; all kept alive by the caller until
 * naja_cond_run_destroy. Between steps the caller may rewrite A0, b0,
 * extra_A and extra_b in place (new bounds of the same conditions, a new
 * request's rows): the next step samples them, provided the row counts stay
 * and a block that was a [T; -T] pair at creation stays one (each packed row
 * holds a row and its negation). p: as naja_chr_plan_create; chains of at
 * most 16384 samples (the LDS transform of the diagnostics). Steps on one run
 * must not overlap.
 */
typedef struct naja_cond_run naja_cond_run;
int naja_cond_run_create(const float* A0, const double* b0, int m0, int d, int n_variants, const int* extra_rows,
                         const float* extra_A, const double* extra_b, float theta, float exchange_tol,
                         const naja_chr_params* p, const float* y0, const float* c, float* P, float* samples,
                         float* ess, float* rhat, float* ess_min, float* rhat_max, int timed, void* stream,
                         naja_cond_run** out);
int naja_cond_run_step(naja_cond_run* run, void* stream);
int naja_cond_run_sampling_ms(const naja_cond_run* run, float* ms);
int naja_cond_run_destroy(naja_cond_run* run);

/*
 * Low-rank whitening P_v = (I + V_v V_v^T)^{-1/2} (expand = 0) or expansion
 * P_v = (I + V_v V_v^T)^{1/2} (expand = 1) from the learned preconditioner's
 * factors V (n_variants x d x rank). No reference counterpart.
 */
int naja_precond_lowrank(const double* V, int n_variants, int d, int rank, int expand, double* P,
                         int nthreads);

/*
 * The learned map P_v = P0_v (I + V_v V_v^T)^{1/2} (expand = 1) or
 * P0_v (I + V_v V_v^T)^{-1/2} (expand = 0) on the GPU, FP64 inside: the
 * r x r matrix function of V^T V by a Newton-Schulz square root, then
 * P = P0 + (P0 V) M V^T. Device memory: P0 and P n_variants x d x d, V
 * n_variants x d x rank, all FP32 row-major; P must not alias P0.
 * rank <= 32 and d <= 160, else NAJA_ERR_UNSUPPORTED. Synchronises the
 * stream (per-variant convergence is checked: NAJA_ERR_RUNTIME).
 * No reference counterpart.
 */
int naja_lowrank_maps_device(const float* P0, const float* V, int n_variants, int d, int rank, int expand,
                             float* P, void* stream);

/* ---------------------------------------------------------------------------
 * Learned preconditioner (no reference counterpart: the reference's
 * preconditioners are the Dikin rounding above and the Jacobi pair schedule,
 * runtime_warmup.cpp:16-108). A 3-layer MLP maps a condition's extra-bound
 * features (n_features) through hidden1, hidden2 ReLU units to a dim x rank
 * factor V = out_scale * o; its whitening is P = (I + V V^T)^{-1/2}
 * (naja_precond_lowrank). Products run on BF16 MFMA with FP32 accumulation;
 * weights are FP32 masters with BF16 copies; loss and gradient in FP64.
 * Parameter vector layout (FP32): W1 (hidden1 x n_features), b1, W2
 * (hidden2 x hidden1), b2, W3 (dim*rank x hidden2), b3 (dim*rank).
 * All data pointers below are device memory; stream may be NULL.
 * ------------------------------------------------------------------------- */
typedef struct naja_lp naja_lp;

/* Covariance models of naja_lp_create's `model`. */
#define NAJA_LP_PRECISION 0  /* precision I + V V^T of S in base-polytope units; whitening (I + V V^T)^{-1/2} */
#define NAJA_LP_COVARIANCE 1 /* S ~ tau (I + V V^T) in a base rounding's frame, tau free; update (I + V V^T)^{1/2} */

/* rank must be 32 and rank <= dim <= 160 (NAJA_ERR_UNSUPPORTED otherwise). */
int naja_lp_create(int n_features, int dim, int rank, int hidden1, int hidden2, int model, uint64_t seed,
                   float out_scale, naja_lp** out);
int naja_lp_destroy(naja_lp* lp);
int naja_lp_param_count(const naja_lp* lp, int64_t* n);
int naja_lp_get_params(const naja_lp* lp, float* dst, void* stream);
/* Loads parameters and resets the Adam state. */
int naja_lp_set_params(naja_lp* lp, const float* src, void* stream);
/* V: batch x dim x rank. */
int naja_lp_forward(naja_lp* lp, const float* features, int batch, float* V, void* stream);
/*
 * Gaussian negative log likelihood of the covariances S (batch x dim x dim)
 * under the handle's model, M = I + V^T V:
 *   precision:  loss[b] = tr(V^T S V) - logdet M;
 *   covariance: loss[b] = dim log(q / dim) + logdet M, q = tr S - tr(M^{-1} V^T S V).
 * grad (nullable) = gradient of the batch mean with respect to the parameters.
 */
int naja_lp_loss_grad(naja_lp* lp, const float* features, const float* S, int batch, float* loss,
                      float* grad, void* stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch-mean loss. */
int naja_lp_train_step(naja_lp* lp, const float* features, const float* S, int batch, float lr,
                       float* loss, void* stream);

/*
 * Per-variant sample covariance of pilot chains: samples n_variants x rows x d
 * (FP32); mean (n_variants x d, FP64); S = scale / (rows - 1) * centred Gram
 * (n_variants x d x d, FP32). FP64 accumulation.
 */
int naja_pilot_covariance(const float* samples, int n_variants, int64_t rows, int d, double scale,
                          float* S, double* mean, void* stream);

/*
 * C = alpha * op(A) op(B), op(A) M x K, op(B) K x N, row-major FP32 storage,
 * operands rounded to BF16 and multiplied on MFMA with FP32 accumulation (the
 * products of the learned preconditioner). trans_a: A stored K x M;
 * trans_b: B stored N x K. rowsum (optional, trans_a only): the M sums over k
 * of op(A) from the unrounded FP32 operand (how the bias gradients are taken).
 * workspace (optional, workspace_floats floats of device memory): lets K be
 * split across workgroups when M x N alone gives too few; the split needs
 * splits * (M * N + M) floats and is skipped when they do not fit.
 */
int naja_gemm_bf16(int trans_a, int trans_b, int M, int N, int K, const float* A, int64_t lda,
                   const float* B, int64_t ldb, float* C, int64_t ldc, float alpha, float* rowsum,
                   float* workspace, int64_t workspace_floats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NAJA_AMD_H */
