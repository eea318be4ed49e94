/*
 * egs_oracle.h -- CPU restatement of ExaGeoStatCPP's maximum-likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * (exageostatcpp_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links, loads or
 * calls anything here.
 *
 * Every function cites the reference file:line it restates.  Matrices are
 * column-major (LAPACK layout) with leading dimension n unless stated.
 */
#ifndef EGS_ORACLE_H
#define EGS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers --------------------------------------------------- */

/* glibc rand()/srand() (TYPE_3 additive feedback generator, the generator the
 * reference reaches through std::srand/rand in ExaGeoStat.cpp:76 and
 * LocationGenerator.cpp:84-89, PredictionHelpers.cpp:148). */
typedef struct {
    int32_t r[34];
    int f, b;           /* front / rear indices into the 31-word ring */
} orc_rand_t;
void orc_srand(orc_rand_t *s, unsigned int seed);
int orc_rand(orc_rand_t *s);

/* LAPACK dlarnv(idist=3, iseed, n, x): N(0,1) by Box-Muller over the dlaruv
 * 48-bit multiplicative congruential stream (call site
 * LinearAlgebraMethods.cpp:348,364).  iseed is updated like LAPACK's. */
void orc_dlarnv_normal(int iseed[4], int64_t n, double *x);
void orc_dlaruv(int iseed[4], int n, double *u); /* n <= 128 uniforms */

/* ---- locations (LocationGenerator.cpp:25-123, ByteHandler.cpp:19-47) ---- */
uint64_t orc_spread_bits(uint64_t v);
uint64_t orc_reverse_spread_bits(uint64_t v);
/* dim: 0 = 2D, 1 = 3D.  z may be NULL for 2D.  Consumes rand() from s. */
void orc_generate_locations(int n, int dim, orc_rand_t *s, double *x, double *y, double *z);
void orc_sort_locations(int n, int dim, double *x, double *y, double *z);

/* ---- special functions -------------------------------------------------- */
/* Modified Bessel function of the second kind K_nu(x), nu >= 0, x > 0
 * (third-party GSL gsl_sf_bessel_Knu at UnivariateMaternStationary.cpp:61-62).
 * Temme's series for x <= 2, Steed's continued fraction CF2 for x > 2,
 * forward recurrence in the order. */
double orc_bessel_knu(double nu, double x);

/* ---- distances (DistanceCalculationHelpers.cpp:21-67) ------------------- */
double orc_distance_earth(double lat1, double lon1, double lat2, double lon2);

/* ---- covariance kernels ------------------------------------------------- */
/* One covariance tile, column-major rows x cols with ld = rows, exactly as
 * Kernel<T>::GenerateCovarianceMatrix (Kernel.hpp:93-97).
 * kernel 0 = UnivariateMaternStationary (UnivariateMaternStationary.cpp:39-66)
 * kernel 1 = BivariateMaternParsimonious (BivariateMaternParsimonious.cpp:40-108)
 * z1/z2 may be NULL (2D).  metric 0 = Euclidean, 1 = great-circle km. */
int orc_covariance_tile(int kernel, double *A, int rows, int cols, int row_off, int col_off,
                        const double *x1, const double *y1, const double *z1,
                        const double *x2, const double *y2, const double *z2,
                        const double *theta, int metric);

/* Full matrix built tile by tile as DCMGCodelet::InsertTask does
 * (dcmg-codelet.cpp:47-74): m x n column-major with ld = m; lower tiles only
 * (row >= col) when lower != 0 and m == n, otherwise every tile. */
int orc_covariance_matrix(int kernel, double *C, int64_t m, int64_t n, int dts, int lower,
                          const double *x1, const double *y1, const double *z1,
                          const double *x2, const double *y2, const double *z2,
                          const double *theta, int metric);

/* ---- dense linear algebra (Chameleon v1.1.0 tile algorithms) ------------ */
/* Right-looking tiled Cholesky, lower, in place (CHAMELEON_dpotrf_Tile call at
 * ChameleonDense.cpp:29).  Returns 0, or k+1 when the leading minor of order
 * k+1 is not positive definite. */
int64_t orc_potrf(double *A, int64_t n, int nb);
/* sum_i log(L_ii) over L_ii > 0 (dmdet-codelet.cpp:70-80). */
double orc_logdet_half(const double *L, int64_t n);
/* z <- L^{-1} z (CHAMELEON_dtrsm_Tile Left/Lower/NoTrans/NonUnit,
 * ChameleonImplementation.cpp:337-347). */
void orc_trsv_lower(const double *L, int64_t n, double *z);
/* z <- L^{-T} z */
void orc_trsv_lower_trans(const double *L, int64_t n, double *z);
/* z <- L z (CHAMELEON_dtrmm_Tile, LinearAlgebraMethods.cpp:400-401). */
void orc_trmv_lower(const double *L, int64_t n, double *z);
double orc_dot(const double *a, const double *b, int64_t n);

/* ---- the hot path --------------------------------------------------------- */
/* One log-likelihood evaluation (ChameleonImplementation.cpp:35-327,
 * univariate/bivariate non-profile branch :257-265).  z (length n_full) is
 * overwritten with L^{-1} z like the reference.  Returns 0 or a non-zero
 * status (not SPD).  Scratch C must hold n_full*n_full doubles. */
int orc_loglik(int kernel, int64_t n_loc, int dts,
               const double *x, const double *y, const double *zc,
               const double *theta, int metric, double *z, double *C,
               double *loglik, double *logdet, double *quad,
               double *t_gen, double *t_facto, double *t_solve);

/* Synthetic observations z = L(theta) e, e = dlarnv(3, {s,s,s,1})
 * (LinearAlgebraMethods.cpp:333-449). */
int orc_generate_observations(int kernel, int64_t n_loc, int dts,
                              const double *x, const double *y, const double *zc,
                              const double *theta, int metric, int seed,
                              double *z_out, double *C);

/* Kriging (LinearAlgebraMethods.cpp:451-605): C22 = Sigma(obs, obs) lower,
 * C12 = Sigma(miss, obs) full, dposv(C22, zobs), zmiss = C12 * x.
 * Returns the MSPE when z_actual != NULL (else -1) through *mspe. */
int orc_krige(int kernel, int64_t n_obs, int64_t n_miss, int dts,
              const double *xo, const double *yo, const double *xm, const double *ym,
              const double *theta, int metric, const double *zobs,
              const double *z_actual, double *zmiss, double *mspe);

/* Prediction errors over tiles of dts values, as the dmse / dmse-bivariate
 * codelets accumulate them (dmse-codelet.cpp:62-75, dmse-bivariate-codelet.cpp:77-100)
 * and LinearAlgebraMethods.cpp:575-600 normalizes them: out[0] = sum of all
 * squared errors / nm; for the bivariate kernel out[1], out[2] = sums over the
 * even (first variable) and odd (second variable) entries / (nm / 2), else 0. */
void orc_mspe(int kernel, int64_t nm, int dts, const double *z_actual, const double *zmiss, double *out);

/* Fisher-Yates shuffle of PredictionHelpers.cpp:141-168 (one array). */
void orc_shuffle(orc_rand_t *s, double *z, double *x, double *y, int n);

#ifdef __cplusplus
}
#endif
#endif
