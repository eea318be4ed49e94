/*
 * exageostat_amd.h -- C ABI of the MI355X (gfx950) maximum-likelihood path.
 *
 * This is the drop-in boundary below ExaGeoStatCPP's dense linear-algebra
 * solver: every entry point names the reference interface it replaces
 * (paths relative to the ExaGeoStatCPP source tree).  Plain pointers and
 * sizes only; host arrays are owned by the caller and copied in; every device
 * buffer is owned by the library.  Functions return EGS_SUCCESS (0) or an
 * error code, and egs_last_error() gives the message of the last failure on
 * the calling thread.  A context is used by one host thread at a time; calls
 * are blocking unless stated otherwise.
 */
#ifndef EXAGEOSTAT_AMD_H
#define EXAGEOSTAT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EGS_SUCCESS 0
#define EGS_ERR_INVALID_ARG 1
#define EGS_ERR_HIP 2
#define EGS_ERR_NOT_SPD 3
#define EGS_ERR_OUT_OF_MEMORY 4
#define EGS_ERR_UNSUPPORTED 5
#define EGS_ERR_NO_DEVICE 6
#define EGS_ERR_STATE 7

/* Covariance kernels (the PluginRegistry names of
 * src/kernels/concrete/UnivariateMaternStationary.cpp:35-36 and
 * src/kernels/concrete/BivariateMaternParsimonious.cpp:35-36). */
#define EGS_KERNEL_UNIVARIATE_MATERN_STATIONARY 0
#define EGS_KERNEL_BIVARIATE_MATERN_PARSIMONIOUS 1

/* Distance metrics (DistanceCalculationHelpers.cpp:38-67): 0 Euclidean,
 * 1 great-circle (haversine, km, x = latitude, y = longitude in degrees). */
#define EGS_METRIC_EUCLIDEAN 0
#define EGS_METRIC_GREAT_CIRCLE 1

/* Precision plans for the factorization.  MIXED: FP64 near the diagonal,
 * FP32 further out, and beyond that FP16 MFMA products of split operands
 * (each panel row scaled by a power of two, hi = fp16(x), lo = fp16(x - hi),
 * hi*hi + hi*lo + lo*hi, FP32 accumulation).  MIXED_BF16: one BF16 product of
 * the unscaled values instead.  Both refine z'Sigma^{-1}z with FP64 residuals. */
#define EGS_PRECISION_FP64 0
#define EGS_PRECISION_MIXED 1
#define EGS_PRECISION_MIXED_BF16 2

typedef struct egs_context_s *egs_context_t;
typedef struct egs_problem_s *egs_problem_t;

/* Message of the last failure on this thread ("" if none). */
const char *egs_last_error(void);
/* Library version as 10000 * major + 100 * minor + patch. */
int egs_version(void);

/* ---- hardware ------------------------------------------------------------
 * Replaces ExaGeoStatHardware::InitHardware / the destructor
 * (src/hardware/ExaGeoStatHardware.cpp:126-212): selects HIP device `device`
 * and creates the library's streams. */
int egs_init(int device, egs_context_t *ctx);
int egs_finalize(egs_context_t ctx);
/* Waits for all work of the context. */
int egs_synchronize(egs_context_t ctx);

/* ---- kernel plugin ---------------------------------------------------------
 * Kernel<T>::GenerateCovarianceMatrix (inst/include/kernels/Kernel.hpp:93-97,
 * e.g. src/kernels/concrete/UnivariateMaternStationary.cpp:39-66): writes the
 * rows x cols column-major block (ld = rows) of Sigma(loc1, loc2) whose first
 * element is (row_off, col_off), evaluated on the GPU.  loc*_z may be NULL
 * (2D).  theta has 3 (univariate) or 6 (bivariate) entries. */
int egs_covariance_tile(egs_context_t ctx, int kernel, double *A, int rows, int cols, int row_off, int col_off,
                        const double *x1, const double *y1, const double *z1, int64_t n1,
                        const double *x2, const double *y2, const double *z2, int64_t n2,
                        const double *theta, int metric);

/* The same tile with no device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The output tile and the
 * location slices it reads (rows / P and cols / P locations) are staged in a
 * static device array of 2^22 doubles (32 MiB) that every GPU the library
 * initializes holds until the process ends: rows * cols + 3 * (ceil(rows / P)
 * + ceil(cols / P)) must fit, else EGS_ERR_INVALID_ARG.  Calls that overlap on one
 * GPU share that array and race; results equal egs_covariance_tile's. */
int egs_covariance_tile_noalloc(egs_context_t ctx, int kernel, double *A, int rows, int cols, int row_off,
                                int col_off, const double *x1, const double *y1, const double *z1, int64_t n1,
                                const double *x2, const double *y2, const double *z2, int64_t n2,
                                const double *theta, int metric);

/* ---- problems (a set of locations and its device-resident Sigma) ---------
 * Replaces LinearAlgebraMethods<T>::InitiateDescriptors
 * (src/linear-algebra-solvers/LinearAlgebraMethods.cpp:35-118): the covariance
 * of n_loc locations (dimension n_loc * P with P = 1 or 2 variables) in lower
 * tiles of `dts` (rounded up to a multiple of 64 on the device).
 * z may be NULL for 2D.  metric is used for every covariance of the problem. */
int egs_problem_create(egs_context_t ctx, int kernel, int64_t n_loc, const double *x, const double *y,
                       const double *z, int metric, int dts, egs_problem_t *problem);
int egs_problem_destroy(egs_problem_t problem);
/* Moves the problem to another context of the same device: its later calls
 * run on that context's streams (the serving loop hands idle regions to
 * whichever worker takes their next request).  No counterpart in the
 * reference. */
int egs_problem_set_context(egs_problem_t problem, egs_context_t ctx);
/* Device bytes the problem holds (each allocation counted with its size
 * class, see egs_memory_cache). */
int egs_problem_device_bytes(egs_problem_t problem, int64_t *bytes);
/* Device bytes a problem of this shape and plan holds once evaluated (its
 * tiles, factor workspaces and vectors; under a mixed plan also the panel
 * copies, the FP64 copy of Sigma and the diagonal-tile inverses), before it
 * exists: the footprint the serving loop admits requests by.  dim is 2 or 3;
 * mode and bands as egs_problem_set_precision.  Prediction buffers (grown on
 * demand, 8 * n_miss * (P + 3) bytes) are not included. */
int egs_problem_estimate_bytes(int kernel, int64_t n_loc, int dim, int dts, int mode, int band_fp64, int band_fp32,
                               int64_t *bytes);
/* Device memory of problems comes in size classes (256-byte granules up to
 * 64 KiB, eight classes per octave above).  Released blocks are kept for the
 * next allocation of the same class while live plus cached bytes stay within
 * limit_bytes, because hipFree waits for the whole device; limit 0 (the
 * default) releases every block at once.  limit_bytes < 0 leaves the limit
 * unchanged.  Reports the cached and live bytes of the current device (either
 * pointer may be NULL).  No counterpart in the reference. */
int egs_memory_cache(int64_t limit_bytes, int64_t *cached_bytes, int64_t *live_bytes);

/* Precision plan: EGS_PRECISION_FP64, or a mixed plan with tiles whose tile
 * distance |i - j| <= band_fp64 in FP64, <= band_fp32 in FP32 and the rest
 * stored in FP32 and updated with 16-bit products accumulated in FP32 (needs a
 * tile size that is a multiple of 192).  The panel solve of FP32-stored rows
 * runs on FP32 MFMA for the univariate kernel and on FP64 MFMA for the
 * bivariate one.  Under a mixed plan the factor's diagonal gives log|Sigma| and
 * z'Sigma^{-1}z comes from conjugate gradients preconditioned by the factor,
 * with residuals b - Sigma x evaluated in FP64 until ||r|| <= 1e-12 ||b||. */
int egs_problem_set_precision(egs_problem_t problem, int mode, int band_fp64, int band_fp32);
/* Under EGS_PRECISION_MIXED, a 192 x 192 block of a tile with tile distance
 * |i - j| > band that would take the split FP16 product (beyond band_fp32;
 * for the univariate kernel also the FP64 tiles beyond the FP64-product band,
 * see egs_problem_set_fp64_product_band) is updated with one FP16 product of
 * the row-scaled values instead of the split three (about 11 significant bits
 * per operand instead of 22) when the panel rows it reads are small:
 * smax(its A rows) * smax(its B rows) <= tau * smax(panel)^2, where smax is the
 * largest power-of-two row scale (2^(e+1) for a largest entry f 2^e).
 * tau = INFINITY takes every such block, tau = 0 none but all-zero ones;
 * band < 0 keeps the split product everywhere.  Default: band 0, tau = 2^-16,
 * which keeps a single-product block's error (about 2^-12 K smax(A) smax(B))
 * 16x under the split product's for the panel's largest block
 * (2^-24 K smax^2).
 * Kept across egs_problem_set_precision calls. */
int egs_problem_set_single_product(egs_problem_t problem, int band, double tau);
/* Under EGS_PRECISION_MIXED, a 192 x 192 block of an FP32 tile beyond the
 * single-product band that passes the single-product test and also has
 * smax(A rows) * smax(B rows) <= tau * smax(panel)^2 takes one product of
 * OCP e4m3 values (3 significant bits) of the row-scaled panel on FP8 MFMA,
 * accumulated in FP32, instead of an FP16 one.  Its error bound, about
 * 2^-4 K smax(A) smax(B), equals the FP16 product's at the single-product
 * threshold when tau = 2^-7 times that threshold.  0 turns it off, and so
 * does a single-product tau of infinity (every block then takes the FP16
 * product without a per-block test).
 * Default: 2^-26 for univariate kernels (2^-10 of the default single-product
 * tau: an 8x smaller bound than the FP16 product's there; at 2^-23 a 100k fit
 * drifted 1.7e-3 from the FP64 estimate), 0 for the bivariate one.  Kept across egs_problem_set_precision
 * calls.  No counterpart in the reference. */
int egs_problem_set_fp8_product(egs_problem_t problem, double tau);
/* Under EGS_PRECISION_MIXED, tiles of the FP64 band (|i - j| <= band_fp64)
 * with |i - j| > band keep FP64 storage but are updated with split FP16
 * products accumulated in FP32 and added in FP64 (for univariate kernels one
 * product per block where egs_problem_set_single_product's test passes).
 * Default 0 for univariate
 * kernels (only the diagonal tiles take FP64 products), 2 for the bivariate
 * kernel.  Negative: every tile of the FP64 band does.  Kept across
 * egs_problem_set_precision calls. */
int egs_problem_set_fp64_product_band(egs_problem_t problem, int band);
/* FP64-plan evaluations (egs_loglik) replayed from a hipGraph: the first call
 * captures the launch sequence (observation copy, covariance, factorization,
 * forward solve), later calls copy the new covariance parameters to the device
 * and replay it, saving the per-launch host cost of small problems.  Results
 * are bit-identical to the direct path.  Any change of plan drops the graph;
 * profiling and mixed plans use the direct path.  Off by default; no
 * counterpart in the reference. */
int egs_problem_set_graph(egs_problem_t problem, int enable);
/* Guard of mixed plans.  log|Sigma| is read off the mixed factor L and the
 * refinement corrects only z'Sigma^{-1}z.  With S = L^{-1} Sigma L^{-T} - I the
 * log-determinant is off by about tr S, and the change of the quadratic form
 * under refinement, q0 - q with q0 = z'(L L^T)^{-1} z, is a one-probe estimate
 * of it (w'S w with w = L^{-1} z); its half over max(|loglik|, 1) estimates the
 * relative log-likelihood error.  An evaluation is repeated under the FP64
 * plan (the mixed plan restored afterwards, unfactored) when that estimate
 * exceeds `tolerance`, when the refinement needs more than max_iterations
 * iterations, when it ends above a relative FP64 residual of 1e-10 (a
 * breakdown: the quadratic form is not recovered and the estimate says
 * nothing), or when the mixed factor is not positive definite (the last two
 * also for egs_predict / egs_predict_mspe).  Defaults:
 * 4 iterations and 1e-5 (the default plan's estimate at n = 100k on the
 * exponential field is of order 1e-7); 0 turns either test off.  No
 * counterpart in the reference (its precision setting is ignored). */
int egs_problem_set_fallback(egs_problem_t problem, int max_iterations, double tolerance);
/* Whether the last egs_loglik fell back to the FP64 plan, how many
 * evaluations of the problem did, and the error estimate of its last mixed
 * evaluation (any pointer may be NULL). */
int egs_last_fallback(egs_problem_t problem, int *fell_back, int64_t *total, double *error_estimate);
/* Iterations and final relative FP64 residual of the last refinement. */
int egs_last_refinement(egs_problem_t problem, int *iterations, double *residual);
/* Block updates (192 x 192) of the tiles beyond the single-product band in the
 * last factorization under EGS_PRECISION_MIXED: those done with the split FP16
 * product and those done with one FP16 product (0 and 0 otherwise). */
int egs_last_product_counts(egs_problem_t problem, int64_t *split_blocks, int64_t *single_blocks);
/* The same with the blocks done with one FP8 product (egs_problem_set_fp8_product). */
int egs_last_product_counts_fp8(egs_problem_t problem, int64_t *split_blocks, int64_t *single_blocks,
                                int64_t *fp8_blocks);

/* Observations Z (length n_loc * P): the measurements descriptor Z and its
 * restore copy Z_COPY of ChameleonImplementation.cpp:113-125. */
int egs_problem_set_observations(egs_problem_t problem, const double *z, int64_t n);
int egs_problem_get_observations(egs_problem_t problem, double *z, int64_t n);

/* ---- the likelihood ------------------------------------------------------
 * ChameleonImplementation<T>::ExaGeoStatMLETile
 * (src/linear-algebra-solvers/concrete/chameleon/ChameleonImplementation.cpp:35-327):
 * Sigma(theta) -> Cholesky -> log|Sigma| -> L^{-1} z -> z' z ->
 * loglik = -0.5 z'Sigma^{-1}z - 0.5 log|Sigma| - (n/2) log(2 pi).
 * logdet and quad may be NULL.  Returns EGS_ERR_NOT_SPD when Sigma(theta) is
 * not positive definite (ChameleonDense.cpp:29-32). */
int egs_loglik(egs_problem_t problem, const double *theta, double *loglik, double *logdet, double *quad);

/* egs_loglik for several problems at once (FP64 plan, one context, one tile
 * size): every step of the tiled Cholesky and of the forward solve is one
 * launch for all problems (problem b at step k while k < its tile count), so
 * the latency-bound diagonal-tile chains of small problems run side by side
 * instead of one after another.  theta of problem b is thetas[6 b ..] (its
 * first 3 or 6 entries).  status[b] receives that problem's result
 * (EGS_SUCCESS, EGS_ERR_NOT_SPD, EGS_ERR_INVALID_ARG for its theta) and
 * loglik[b], logdet[b], quad[b] its values (logdet and quad may be NULL).
 * Each problem keeps its factor, as after egs_loglik.  Returns an error only
 * for an invalid batch.  No counterpart in the reference (it evaluates one
 * problem per call). */
int egs_loglik_batch(egs_problem_t *problems, const double *thetas, int count, double *loglik, double *logdet,
                     double *quad, int *status);

/* The phases of egs_loglik, for callers that drive them one by one:
 *   RuntimeFunctions<T>::CovarianceMatrix (src/runtime/starpu/StarPuFunctions.cpp:24-45),
 *   ExaGeoStatPotrfTile (ChameleonDense.cpp:25-37),
 *   ExaGeoStatMeasureDetTileAsync + ExaGeoStatTrsmTile + ExaGeoStatDoubleDotProduct
 *   (StarPuFunctions.cpp:119-135, 189-204; ChameleonImplementation.cpp:337-347).
 * egs_potrf factors the matrix of the last egs_covariance call and fails with
 * EGS_ERR_STATE when there is none (never generated, or already factored);
 * egs_solve_logdet_quad needs a successful egs_potrf (or evaluation). */
int egs_covariance(egs_problem_t problem, const double *theta);
int egs_potrf(egs_problem_t problem);
int egs_solve_logdet_quad(egs_problem_t problem, double *logdet, double *quad);

/* ChameleonImplementation<T>::ExaGeoStatTrsmTile with side Left, Lower,
 * NonUnit (ChameleonImplementation.cpp:337-347): z <- alpha L^{-1} z
 * (trans 0) or alpha L^{-T} z (trans 1) with the factor of the last egs_potrf
 * or evaluation, for a host vector of n_loc * P entries (the reference's Z
 * aliases caller memory, LinearAlgebraMethods.cpp:75-77).  FP64 plan only. */
int egs_trsm_vector(egs_problem_t problem, int trans, double alpha, double *z, int64_t n);

/* Copies the lower tiles of the current device matrix (Sigma after
 * egs_covariance, L after egs_potrf) into a column-major n x n host array
 * (ExaGeoStatDesc2Lap, LinearAlgebraMethods.cpp:1373-1380).  The strictly
 * upper part of `out` is left untouched. */
int egs_get_matrix(egs_problem_t problem, double *out, int64_t ld);

/* Phase times of the last egs_loglik, milliseconds: covariance, factorization,
 * solve (log-det, triangular solve and dot product). */
int egs_last_timings(egs_problem_t problem, double *t_covariance, double *t_factor, double *t_solve);

/* ---- measurement -----------------------------------------------------------
 * The reference times phases with START_TIMING/STOP_TIMING
 * (inst/include/utilities/Logger.hpp:153-160; ChameleonImplementation.cpp:129-219).
 * egs_problem_profile(p, 1) resets the counters and brackets every launch of the
 * kernels below with HIP events on the library stream; egs_problem_kernel_stats
 * returns, per kernel, the number of timed launches, their summed duration and
 * their summed algorithmic flops (FP64 real rows only, padding excluded). */
#define EGS_PROFILE_UPDATE 0        /* trailing SYRK/GEMM update, one launch per step */
#define EGS_PROFILE_COVARIANCE 1    /* covariance generation of all lower tiles */
#define EGS_PROFILE_PANEL_TRSM 2    /* panel triangular solve, one launch per step */
#define EGS_PROFILE_DIAGONAL 3      /* diagonal-tile factorization (all its launches) */
#define EGS_PROFILE_SOLVE 4         /* forward solve with log-det and dot product */
#define EGS_PROFILE_UPDATE_F32 5    /* mixed plans: FP32-band update */
#define EGS_PROFILE_UPDATE_16 6      /* mixed plans: update with 16-bit products */
#define EGS_PROFILE_REFINE 7        /* mixed plans: refinement of z'Sigma^{-1}z */
#define EGS_PROFILE_UPDATE_16_SELECT 8  /* mixed plans: update of the tiles beyond the single-product band */
#define EGS_PROFILE_KERNELS 9
int egs_problem_profile(egs_problem_t problem, int enable);
int egs_problem_kernel_stats(egs_problem_t problem, int kernel, int64_t *launches, double *total_ms, double *flops);

/* ---- synthetic data ------------------------------------------------------
 * LocationGenerator<T>::GenerateLocations (src/data-generators/LocationGenerator.cpp:25-123):
 * jittered grid from the C library's rand() (seed it with srand first, as
 * ExaGeoStat.cpp:76 does), then Morton order on a 16-bit grid.
 * dim: 2 or 3 (z must be non-NULL for 3). */
int egs_generate_locations(int64_t n, int dim, double *x, double *y, double *z);

/* LinearAlgebraMethods<T>::GenerateObservationsVector
 * (LinearAlgebraMethods.cpp:333-449): z = L(theta) e with e = dlarnv(3,
 * {seed, seed, seed, 1}); the result is written to z_out (length n_loc * P, may
 * be NULL) and becomes the problem's observations.  L is the FP64 factor under
 * every plan (a mixed plan is restored afterwards, unfactored). */
int egs_generate_observations(egs_problem_t problem, const double *theta, int seed, double *z_out);

/* ---- kriging ---------------------------------------------------------------
 * LinearAlgebraMethods<T>::ExaGeoStatMLEPredictTile (LinearAlgebraMethods.cpp:451-605):
 * zmiss = Sigma(miss, obs) Sigma(obs, obs)^{-1} zobs for the problem's
 * locations and observations.  The factor of Sigma(obs, obs) at theta is
 * reused when the problem already holds it.  z_actual may be NULL; otherwise
 * *mspe receives mean((z_actual - zmiss)^2) (dmse-codelet.cpp:40-73), else -1. */
int egs_predict(egs_problem_t problem, const double *theta, int64_t n_miss, const double *xm, const double *ym,
                const double *zm, double *zmiss, const double *z_actual, double *mspe);
/* The same prediction returning the three errors ExaGeoStatMLEPredictTile
 * returns (LinearAlgebraMethods.cpp:575-600): mspe3[0] over all n_miss * P
 * values, and for the bivariate kernel mspe3[1], mspe3[2] over the first and
 * second variable (the even and odd entries of the interleaved vector, as the
 * dmse-bivariate codelet splits them, dmse-bivariate-codelet.cpp:77-100), each
 * divided by n_miss; 0 and 0 for univariate kernels.  Without z_actual:
 * {-1, 0, 0}. */
int egs_predict_mspe(egs_problem_t problem, const double *theta, int64_t n_miss, const double *xm, const double *ym,
                     const double *zm, double *zmiss, const double *z_actual, double *mspe3);

/* ---- maximum-likelihood driver ----------------------------------------------
 * Bound-constrained derivative-free minimisation (Powell's BOBYQA), the
 * optimiser that StarpuRuntimeSolver<T>::ModelingOperations takes from NLopt
 * (src/runtime-solver/concrete/StarpuRuntimeSolver.cpp:38-48).  fn is called
 * with a point inside [lb, ub]; x holds the start on entry and the best point
 * on return.  Stops when an improvement of the best value is below ftol_abs
 * (0: off), after maxeval evaluations, or when the trust region shrinks to
 * rho_end times the initial step (<= 0: 1e-7).  *reason: 1 radius, 2 ftol,
 * 3 maxeval.  A NaN from fn ends the search with EGS_ERR_INVALID_ARG. */
typedef double (*egs_objective_fn)(const double *x, void *user);
int egs_bobyqa(int n, double *x, const double *lb, const double *ub, double ftol_abs, int maxeval, double rho_end,
               egs_objective_fn fn, void *user, double *fmin, int *nevals, int *reason);

/* StarpuRuntimeSolver<T>::ModelingOperations (StarpuRuntimeSolver.cpp:28-65):
 * maximises egs_loglik over theta in [lb, ub] from start (NULL: lb, as
 * Configurations::InitializeAllTheta sets it, Configurations.cpp:200) with
 * ftol_abs = tolerance and at most max_iterations likelihood evaluations.
 * theta receives the best point, *loglik the best value (NLopt's
 * last_optimum_value), *nevals the evaluation count (may be NULL).  A failed
 * evaluation (not positive definite, invalid theta) ends the search with its
 * status, as the reference's exception does.  Under a mixed plan the search
 * ends with an FP64 polish within the same evaluation budget: a second search
 * under the FP64 plan from the mixed optimum (first steps 0.05 (ub - lb)),
 * whose optimum is the one returned (the mixed plan is restored). */
int egs_data_modeling(egs_problem_t problem, const double *lb, const double *ub, const double *start, double tolerance,
                      int max_iterations, double *theta, double *loglik, int *nevals);

/* ---- serving ------------------------------------------------------------------
 * A request loop over many spatial regions on one GPU.  The reference has no
 * serving layer: it evaluates one problem per process through
 * ExaGeoStatDataModeling / ExaGeoStatPrediction (ExaGeoStat.cpp:91-125); the
 * server runs those two operations (egs_loglik, egs_predict) for registered
 * regions concurrently.  Requests wait in one queue; each of `workers`
 * contexts (1..16, each with its own pair of streams) takes the oldest request
 * whose region is idle, so a region's requests run in submission order.  A
 * worker that takes a likelihood request of a small FP64 region (at most 16
 * tiles) also takes the other queued ones of the same tile size (up to 32)
 * and evaluates them together with egs_loglik_batch.  A region stays resident
 * after its request: a prediction at the theta of the region's last
 * evaluation reuses the factor kept in HBM.  Regions are admitted by their
 * footprint (egs_problem_estimate_bytes) within hbm_budget_bytes, releasing
 * the least recently used idle ones; while the server lives, released device
 * blocks are cached for later regions within the same budget
 * (egs_memory_cache, set back to 0 by egs_server_destroy).  A mixed-plan
 * region whose tiles all lie in the FP64 band runs the FP64 plan; one whose
 * evaluation falls back to FP64 (egs_problem_set_fallback) keeps the FP64
 * plan for its later requests. */
typedef struct egs_server_s *egs_server_t;
int egs_server_create(int device, int64_t hbm_budget_bytes, int workers, egs_server_t *server);
int egs_server_destroy(egs_server_t server);
/* Copies the locations and observations (n_loc * P values); *region is its
 * id.  The region is evaluated under the FP64 plan. */
int egs_server_add_region(egs_server_t server, int kernel, int64_t n_loc, const double *x, const double *y,
                          const double *z, const double *obs, int metric, int dts, int *region);
/* The same with a precision plan (mode and bands as egs_problem_set_precision;
 * a mixed plan needs a tile size that is a multiple of 192).  Fails with
 * EGS_ERR_OUT_OF_MEMORY when the region's footprint exceeds the budget. */
int egs_server_add_region_plan(egs_server_t server, int kernel, int64_t n_loc, const double *x, const double *y,
                               const double *z, const double *obs, int metric, int dts, int mode, int band_fp64,
                               int band_fp32, int *region);
int egs_server_submit_loglik(egs_server_t server, int region, const double *theta, int64_t *ticket);
int egs_server_submit_predict(egs_server_t server, int region, const double *theta, int64_t n_miss, const double *xm,
                              const double *ym, int64_t *ticket);
/* Blocks until the request is done; *value is the log-likelihood (or -1 for a
 * prediction), zmiss (n_miss * P values) the predictions, *latency_ms the time
 * from submission to completion.  A failed request returns its error. */
int egs_server_wait(egs_server_t server, int64_t ticket, double *value, double *zmiss, int64_t zmiss_len,
                    double *latency_ms);
/* Completed requests, latency percentiles (ms), completions per second since
 * the first submission, and the HBM held by resident regions. */
int egs_server_stats(egs_server_t server, int64_t *completed, double *p50_ms, double *p99_ms, double *per_second,
                     int64_t *resident_bytes);
/* Batches run and the requests they held; predictions, and those that found
 * their region's factor at their theta still in HBM; mixed-plan regions that
 * switched to the FP64 plan after an evaluation fell back to it. */
int egs_server_batch_stats(egs_server_t server, int64_t *batches, int64_t *batched_requests, int64_t *predicts,
                           int64_t *predict_hits, int64_t *fp64_regions);
int egs_server_reset_stats(egs_server_t server);

#ifdef __cplusplus
}
#endif
#endif
