/*
 * egs_oracle.c -- CPU restatement of ExaGeoStatCPP's maximum-likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Each function names the reference file:line it follows.  Third-party
 * arithmetic the reference pulls from un-vendored libraries is restated from
 * its published algorithm:
 *   - glibc 2.x random_r TYPE_3 (rand/srand),
 *   - LAPACK 3.x dlaruv/dlarnv (OpenBLAS 0.3.21 / MKL in the reference build),
 *   - GSL 2.7.1 gsl_sf_bessel_Knu (Temme 1975 series; Steed's CF2 as in
 *     Thompson & Barnett 1987),
 *   - Chameleon v1.1.0 dpotrf/dtrsm/dposv/dgemm tile algorithms.
 */
#include "egs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793      /* Definitions.hpp:45 */
#define ORC_EARTH_RADIUS 6371.0       /* Definitions.hpp:50 */

static double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

/* ======================================================================== */
/* glibc rand(): r[i] = r[i-3] + r[i-31] (mod 2^32) after a Lehmer seeding of  */
/* 31 words and 310 discarded outputs; the output is the sum shifted right.  */
/* ======================================================================== */
void orc_srand(orc_rand_t *s, unsigned int seed) {
    int32_t *r = s->r;
    int i;
    r[0] = (int32_t) (seed == 0 ? 1u : seed);
    for (i = 1; i < 31; i++) {
        /* 16807 * r mod (2^31 - 1) by Schrage's method */
        int32_t hi = r[i - 1] / 127773;
        int32_t lo = r[i - 1] % 127773;
        int32_t w = 16807 * lo - 2836 * hi;
        if (w < 0) w += 2147483647;
        r[i] = w;
    }
    s->f = 3;
    s->b = 0;
    for (i = 0; i < 310; i++) (void) orc_rand(s);
}

int orc_rand(orc_rand_t *s) {
    uint32_t v = (uint32_t) s->r[s->f] + (uint32_t) s->r[s->b];
    s->r[s->f] = (int32_t) v;
    int out = (int) (v >> 1);
    s->f = (s->f + 1) % 31;
    s->b = (s->b + 1) % 31;
    return out;
}

/* ======================================================================== */
/* LAPACK dlaruv / dlarnv(idist = 3).                                         */
/* dlaruv returns x_i = (seed * a^i mod 2^48) / 2^48 for i = 1..n with        */
/* a = 33952834046453 and leaves seed * a^n as the new seed, so successive  */
/* calls walk one multiplicative congruential sequence.                      */
/* ======================================================================== */
static const uint64_t ORC_LCG_A = 33952834046453ull;
static const uint64_t ORC_MASK48 = (1ull << 48) - 1;

void orc_dlaruv(int iseed[4], int n, double *u) {
    uint64_t s = ((((uint64_t) iseed[0] * 4096u + (uint64_t) iseed[1]) * 4096u + (uint64_t) iseed[2]) * 4096u
                  + (uint64_t) iseed[3]) & ORC_MASK48;
    for (int i = 0; i < n; i++) {
        s = (s * ORC_LCG_A) & ORC_MASK48;
        u[i] = (double) s * 0x1p-48;
    }
    iseed[0] = (int) ((s >> 36) & 4095u);
    iseed[1] = (int) ((s >> 24) & 4095u);
    iseed[2] = (int) ((s >> 12) & 4095u);
    iseed[3] = (int) (s & 4095u);
}

void orc_dlarnv_normal(int iseed[4], int64_t n, double *x) {
    double u[128];
    const double two_pi = 6.2831853071795864769252867663;
    for (int64_t iv = 0; iv < n; iv += 64) {
        int il = (int) ((n - iv) < 64 ? (n - iv) : 64);
        orc_dlaruv(iseed, 2 * il, u);
        for (int i = 0; i < il; i++)
            x[iv + i] = sqrt(-2.0 * log(u[2 * i])) * cos(two_pi * u[2 * i + 1]);
    }
}

/* ======================================================================== */
/* Morton order (ByteHandler.cpp:19-47) and location generation              */
/* (LocationGenerator.cpp:25-123).                                           */
/* ======================================================================== */
uint64_t orc_spread_bits(uint64_t v) {
    /* bit b of the 16-bit input moves to bit 4b */
    uint64_t out = 0;
    for (int b = 0; b < 16; b++) out |= ((v >> b) & 1ull) << (4 * b);
    return out;
}

uint64_t orc_reverse_spread_bits(uint64_t v) {
    uint64_t out = 0;
    for (int b = 0; b < 16; b++) out |= ((v >> (4 * b)) & 1ull) << b;
    return out;
}

static double orc_uniform(orc_rand_t *s, double lo, double hi) {
    double r = (double) orc_rand(s) / (1.0 + 2147483647.0);
    return r * (hi - lo) + lo;
}

static int orc_cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *) a, y = *(const uint64_t *) b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

void orc_sort_locations(int n, int dim, double *x, double *y, double *z) {
    uint64_t *code = (uint64_t *) malloc(sizeof(uint64_t) * (size_t) n);
    for (int i = 0; i < n; i++) {
        uint16_t qx = (uint16_t) (x[i] * 65535.0 + .5);
        uint16_t qy = (uint16_t) (y[i] * 65535.0 + .5);
        uint16_t qz = (dim != 0 && z) ? (uint16_t) (z[i] * 65535.0 + .5) : 0;
        code[i] = (orc_spread_bits(qz) << 2) + (orc_spread_bits(qy) << 1) + orc_spread_bits(qx);
    }
    qsort(code, (size_t) n, sizeof(uint64_t), orc_cmp_u64);
    for (int i = 0; i < n; i++) {
        x[i] = (double) orc_reverse_spread_bits(code[i]) / 65535.0;
        y[i] = (double) orc_reverse_spread_bits(code[i] >> 1) / 65535.0;
        if (dim != 0 && z) z[i] = (double) orc_reverse_spread_bits(code[i] >> 2) / 65535.0;
    }
    free(code);
}

void orc_generate_locations(int n, int dim, orc_rand_t *s, double *x, double *y, double *z) {
    int root = dim != 0 ? (int) ceil(cbrt((double) n)) : (int) ceil(sqrt((double) n));
    int idx = 0;
    for (int i = 0; i < root && idx < n; i++) {
        for (int j = 0; j < root && idx < n; j++) {
            if (dim != 0) {
                for (int k = 0; k < root && idx < n; k++) {
                    x[idx] = ((i + 1) - 0.5 + orc_uniform(s, -0.4, 0.4)) / root;
                    y[idx] = ((j + 1) - 0.5 + orc_uniform(s, -0.4, 0.4)) / root;
                    z[idx] = ((k + 1) - 0.5 + orc_uniform(s, -0.4, 0.4)) / root;
                    idx++;
                }
            } else {
                x[idx] = ((i + 1) - 0.5 + orc_uniform(s, -0.4, 0.4)) / root;
                y[idx] = ((j + 1) - 0.5 + orc_uniform(s, -0.4, 0.4)) / root;
                idx++;
            }
        }
    }
    orc_sort_locations(n, dim, x, y, z);
}

/* ======================================================================== */
/* K_nu(x).  With nu = N + mu, |mu| <= 1/2, compute K_mu and K_mu+1, then    */
/* K_{v+1} = K_{v-1} + (2v/x) K_v upward.                                    */
/*  x <= 2 : Temme's series.  Gamma1, Gamma2 come from the Taylor series of  */
/*           1/Gamma(1+t) = sum a_k t^k (coefficients evaluated to 25 digits   */
/*           with mpmath; a_1 is Euler's constant).                         */
/*  x >  2 : Steed's evaluation of the continued fraction CF2 together with  */
/*           the companion sum that normalises it (Thompson-Barnett).        */
/* ======================================================================== */
static const double orc_rgam_taylor[] = {
    1.0, 0.5772156649015328606065121, -0.6558780715202538810770195, -0.04200263503409523552900393,
    0.1665386113822914895017008, -0.0421977345555443367482083, -0.009621971527876973562114922,
    0.00721894324666309954239501, -0.001165167591859065112113971, -0.00021524167411495097281573,
    0.0001280502823881161861531986, -0.00002013485478078823865568939, -0.000001250493482142670657345359,
    0.00000113302723198169588237413, -0.0000002056338416977607103450154, 6.116095104481415817862499e-9,
    5.002007644469222930055665e-9, -1.181274570487020144588127e-9, 1.04342671169110051049154e-10,
    7.782263439905071254049937e-12, -3.696805618642205708187816e-12, 5.100370287454475979015481e-13,
    -2.05832605356650678322243e-14, -5.348122539423017982370017e-15, 1.226778628238260790158894e-15,
    -1.181259301697458769513765e-16, 1.186692254751600332579777e-18, 1.412380655318031781555804e-18,
    -2.298745684435370206592479e-19, 1.714406321927337433383963e-20, 1.337351730493693114864781e-22,
    -2.054233551766672789325025e-22, 2.73603004860799984483151e-23};

/* gamma1 = (1/G(1-m) - 1/G(1+m)) / (2m), gamma2 = (1/G(1-m) + 1/G(1+m)) / 2 */
static void orc_temme_gammas(double mu, double *g1, double *g2) {
    double m2 = mu * mu;
    double odd = 0.0, even = 0.0;
    for (int k = 31; k >= 1; k -= 2) odd = odd * m2 + orc_rgam_taylor[k];
    for (int k = 32; k >= 0; k -= 2) even = even * m2 + orc_rgam_taylor[k];
    *g1 = -odd;
    *g2 = even;
}

static void orc_k_temme(double mu, double x, double *kmu, double *kmu1) {
    const double eps = 1e-17;
    double half_x = 0.5 * x;
    double lg = -log(half_x);            /* ln(2/x) */
    double sig = mu * lg;
    double fact = (mu == 0.0) ? 1.0 : (ORC_PI * mu) / sin(ORC_PI * mu);
    double shc = (sig == 0.0) ? 1.0 : sinh(sig) / sig;
    double g1, g2;
    orc_temme_gammas(mu, &g1, &g2);
    double rg_plus = g2 - mu * g1;       /* 1/Gamma(1+mu) */
    double rg_minus = g2 + mu * g1;      /* 1/Gamma(1-mu) */
    double f = fact * (g1 * cosh(sig) + g2 * shc * lg);
    double es = exp(sig);                /* (2/x)^mu */
    double p = 0.5 * es / rg_plus;       /* (1/2) (x/2)^-mu Gamma(1+mu) */
    double q = 0.5 / (es * rg_minus);    /* (1/2) (x/2)^mu  Gamma(1-mu) */
    double c = 1.0;
    double quarter_x2 = half_x * half_x;
    double s0 = f, s1 = p;
    for (int k = 1; k < 500; k++) {
        double dk = (double) k;
        f = (dk * f + p + q) / (dk * dk - mu * mu);
        c *= quarter_x2 / dk;
        p /= (dk - mu);
        q /= (dk + mu);
        double t0 = c * f;
        s0 += t0;
        s1 += c * (p - dk * f);
        if (fabs(t0) < eps * fabs(s0)) break;
    }
    *kmu = s0;
    *kmu1 = s1 * (2.0 / x);
}

static void orc_k_cf2(double mu, double x, double *kmu, double *kmu1) {
    const double eps = 1e-17;
    double b = 2.0 * (1.0 + x);
    double d = 1.0 / b;
    double h = d, dh = d;
    double q1 = 0.0, q2 = 1.0;
    double a1 = 0.25 - mu * mu;
    double q = a1, c = a1;
    double a = -a1;
    double s = 1.0 + q * dh;
    for (int i = 2; i < 100000; i++) {
        a -= 2.0 * (i - 1);
        c = -a * c / i;
        double qn = (q1 - b * q2) / a;
        q1 = q2;
        q2 = qn;
        q += c * qn;
        b += 2.0;
        d = 1.0 / (b + a * d);
        dh = (b * d - 1.0) * dh;
        h += dh;
        double ds = q * dh;
        s += ds;
        if (fabs(ds) < eps * fabs(s)) break;
    }
    h *= a1;
    double k0 = sqrt(ORC_PI / (2.0 * x)) * exp(-x) / s;
    *kmu = k0;
    *kmu1 = k0 * (mu + x + 0.5 - h) / x;
}

double orc_bessel_knu(double nu, double x) {
    if (!(x > 0.0) || nu < 0.0) return NAN;
    int nl = (int) floor(nu + 0.5);
    double mu = nu - nl;
    double k0, k1;
    if (x <= 2.0) orc_k_temme(mu, x, &k0, &k1);
    else orc_k_cf2(mu, x, &k0, &k1);
    for (int i = 1; i <= nl; i++) {
        double k2 = (mu + i) * (2.0 / x) * k1 + k0;
        k0 = k1;
        k1 = k2;
    }
    return k0;
}

/* ======================================================================== */
/* Distances (DistanceCalculationHelpers.cpp:21-67).                         */
/* ======================================================================== */
double orc_distance_earth(double lat1, double lon1, double lat2, double lon2) {
    double lat1r = lat1 * ORC_PI / 180, lon1r = lon1 * ORC_PI / 180;
    double lat2r = lat2 * ORC_PI / 180, lon2r = lon2 * ORC_PI / 180;
    double u = sin((lat2r - lat1r) / 2);
    double v = sin((lon2r - lon1r) / 2);
    return 2.0 * ORC_EARTH_RADIUS * asin(sqrt(u * u + cos(lat1r) * cos(lat2r) * v * v));
}

static double orc_distance(const double *x1, const double *y1, const double *z1,
                           const double *x2, const double *y2, const double *z2,
                           int i, int j, int metric, int use_z) {
    double dx = x2[j] - x1[i];
    double dy = y2[j] - y1[i];
    if (!z1 || !z2 || !use_z) {
        if (metric == 1) return orc_distance_earth(x1[i], y1[i], x2[j], y2[j]);
        return sqrt(dx * dx + dy * dy);
    }
    double dz = z2[j] - z1[i];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* ======================================================================== */
/* Covariance kernels.                                                       */
/* ======================================================================== */
static void orc_matern_uni(double *A, int rows, int cols, int row_off, int col_off,
                           const double *x1, const double *y1, const double *z1,
                           const double *x2, const double *y2, const double *z2,
                           const double *theta, int metric) {
    /* UnivariateMaternStationary.cpp:47-65 */
    const double sigma2 = theta[0], nu = theta[2];
    const double con = sigma2 * (1.0 / (pow(2, nu - 1) * tgamma(nu)));
    int use_z = z1 != NULL;
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < cols; j++) {
            double d = orc_distance(x1, y1, z1, x2, y2, z2, row_off + i, col_off + j, metric, use_z) / theta[1];
            A[i + (int64_t) j * rows] = (d == 0.0) ? sigma2 : con * pow(d, nu) * orc_bessel_knu(nu, d);
        }
    }
}

static void orc_matern_bivariate(double *A, int rows, int cols, int row_off, int col_off,
                                 const double *x1, const double *y1, const double *x2, const double *y2,
                                 const double *theta, int metric) {
    /* BivariateMaternParsimonious.cpp:50-107 */
    double con1 = theta[0] / (pow(2, theta[3] - 1) * tgamma(theta[3]));
    double con2 = theta[1] / (pow(2, theta[4] - 1) * tgamma(theta[4]));
    double nu12 = 0.5 * (theta[3] + theta[4]);
    double rho = theta[5] * sqrt((tgamma(theta[3] + 1) * tgamma(theta[4] + 1)) / (tgamma(theta[3]) * tgamma(theta[4])))
                 * tgamma(nu12) / tgamma(nu12 + 1);
    double con12 = rho * sqrt(theta[0] * theta[1]) / (pow(2, nu12 - 1) * tgamma(nu12));
    int64_t total = (int64_t) rows * cols;
    int i0 = row_off / 2;
    for (int i = 0; i < rows; i += 2, i0++) {
        int j0 = col_off / 2;
        for (int j = 0; j < cols; j += 2, j0++) {
            double e = orc_distance(x1, y1, NULL, x2, y2, NULL, i0, j0, metric, 0) / theta[2];
            int64_t p00 = i + (int64_t) j * rows, p10 = (i + 1) + (int64_t) j * rows;
            int64_t p01 = i + (int64_t) (j + 1) * rows, p11 = (i + 1) + (int64_t) (j + 1) * rows;
            if (e == 0) {
                A[p00] = theta[0];
                if (p10 < total) A[p10] = rho * sqrt(theta[0] * theta[1]);
                if (p01 < total) A[p01] = rho * sqrt(theta[0] * theta[1]);
                if (p11 < total) A[p11] = theta[1];
            } else {
                A[p00] = con1 * pow(e, theta[3]) * orc_bessel_knu(theta[3], e);
                if (p10 < total) A[p10] = con12 * pow(e, nu12) * orc_bessel_knu(nu12, e);
                if (p01 < total) A[p01] = con12 * pow(e, nu12) * orc_bessel_knu(nu12, e);
                if (p11 < total) A[p11] = con2 * pow(e, theta[4]) * orc_bessel_knu(theta[4], e);
            }
        }
    }
}

int orc_covariance_tile(int kernel, double *A, int rows, int cols, int row_off, int col_off,
                        const double *x1, const double *y1, const double *z1,
                        const double *x2, const double *y2, const double *z2,
                        const double *theta, int metric) {
    if (metric == 1 && z1 && z2) return -2; /* great-circle distance is 2D only (:60-62) */
    if (kernel == 0) {
        orc_matern_uni(A, rows, cols, row_off, col_off, x1, y1, z1, x2, y2, z2, theta, metric);
        return 0;
    }
    if (kernel == 1) {
        orc_matern_bivariate(A, rows, cols, row_off, col_off, x1, y1, x2, y2, theta, metric);
        return 0;
    }
    return -1;
}

int orc_covariance_matrix(int kernel, double *C, int64_t m, int64_t n, int dts, int lower,
                          const double *x1, const double *y1, const double *z1,
                          const double *x2, const double *y2, const double *z2,
                          const double *theta, int metric) {
    int64_t mt = (m + dts - 1) / dts, nt = (n + dts - 1) / dts;
    int status = 0;
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int64_t col = 0; col < nt; col++) {
        for (int64_t row = 0; row < mt; row++) {
            if (lower && m == n && row < col) continue;
            int rows = (int) (row == mt - 1 ? m - row * dts : dts);
            int cols = (int) (col == nt - 1 ? n - col * dts : dts);
            double *tile = (double *) malloc(sizeof(double) * (size_t) rows * (size_t) cols);
            int st = orc_covariance_tile(kernel, tile, rows, cols, (int) (row * dts), (int) (col * dts),
                                         x1, y1, z1, x2, y2, z2, theta, metric);
            if (st) {
#pragma omp atomic write
                status = st;
            }
            for (int j = 0; j < cols; j++)
                memcpy(C + row * dts + (col * dts + j) * m, tile + (int64_t) j * rows, sizeof(double) * (size_t) rows);
            free(tile);
        }
    }
    return status;
}

/* ======================================================================== */
/* Tiled Cholesky (Chameleon dpotrf, lower, right-looking): for each k,     */
/* POTRF(k,k); TRSM(m,k) for m > k; SYRK(m,m) and GEMM(m,n) for the         */
/* trailing tiles.                                                           */
/* ======================================================================== */
static int64_t orc_potrf_tile(double *A, int64_t ld, int nb) {
    for (int j = 0; j < nb; j++) {
        double d = A[j + j * ld];
        for (int l = 0; l < j; l++) d -= A[j + l * ld] * A[j + l * ld];
        if (!(d > 0.0)) return j + 1;
        d = sqrt(d);
        A[j + j * ld] = d;
        for (int l = 0; l < j; l++) {
            double b = A[j + l * ld];
            for (int i = j + 1; i < nb; i++) A[i + j * ld] -= A[i + l * ld] * b;
        }
        for (int i = j + 1; i < nb; i++) A[i + j * ld] /= d;
    }
    return 0;
}

/* B (rows x nb) <- B L^{-T} with L the nb x nb lower tile. */
static void orc_trsm_tile(const double *L, double *B, int64_t ld, int rows, int nb) {
    for (int c = 0; c < nb; c++) {
        double *bc = B + c * ld;
        for (int l = 0; l < c; l++) {
            double f = L[c + l * ld];
            const double *bl = B + l * ld;
            for (int i = 0; i < rows; i++) bc[i] -= bl[i] * f;
        }
        double inv = 1.0 / L[c + c * ld];
        for (int i = 0; i < rows; i++) bc[i] *= inv;
    }
}

/* C (rows x cols) -= A (rows x k) B^T (B is cols x k); lower_only for SYRK. */
static void orc_gemm_nt(double *C, const double *A, const double *B, int64_t ld, int rows, int cols, int k,
                        int lower_only) {
    for (int j = 0; j < cols; j++) {
        double *cj = C + j * ld;
        int i0 = lower_only ? j : 0;
        for (int l = 0; l < k; l++) {
            double b = B[j + l * ld];
            const double *al = A + l * ld;
            for (int i = i0; i < rows; i++) cj[i] -= al[i] * b;
        }
    }
}

int64_t orc_potrf(double *A, int64_t n, int nb) {
    if (nb <= 0) nb = 256;
    int64_t nt = (n + nb - 1) / nb;
    for (int64_t k = 0; k < nt; k++) {
        int kb = (int) (k == nt - 1 ? n - k * nb : nb);
        double *Akk = A + k * nb + k * nb * n;
        int64_t info = orc_potrf_tile(Akk, n, kb);
        if (info) return k * nb + info;
#pragma omp parallel for schedule(dynamic)
        for (int64_t m = k + 1; m < nt; m++) {
            int mb = (int) (m == nt - 1 ? n - m * nb : nb);
            orc_trsm_tile(Akk, A + m * nb + k * nb * n, n, mb, kb);
        }
#pragma omp parallel for schedule(dynamic) collapse(2)
        for (int64_t jn = k + 1; jn < nt; jn++) {
            for (int64_t m = k + 1; m < nt; m++) {
                if (m < jn) continue;
                int mb = (int) (m == nt - 1 ? n - m * nb : nb);
                int jb = (int) (jn == nt - 1 ? n - jn * nb : nb);
                orc_gemm_nt(A + m * nb + jn * nb * n, A + m * nb + k * nb * n, A + jn * nb + k * nb * n, n, mb, jb,
                            kb, m == jn);
            }
        }
    }
    return 0;
}

double orc_logdet_half(const double *L, int64_t n) {
    double r = 0.0;
    for (int64_t i = 0; i < n; i++)
        if (L[i + i * n] > 0) r += log(L[i + i * n]);
    return r;
}

void orc_trsv_lower(const double *L, int64_t n, double *z) {
    for (int64_t j = 0; j < n; j++) {
        z[j] /= L[j + j * n];
        double zj = z[j];
        const double *lj = L + j * n;
        for (int64_t i = j + 1; i < n; i++) z[i] -= lj[i] * zj;
    }
}

void orc_trsv_lower_trans(const double *L, int64_t n, double *z) {
    for (int64_t j = n - 1; j >= 0; j--) {
        const double *lj = L + j * n;
        double s = z[j];
        for (int64_t i = j + 1; i < n; i++) s -= lj[i] * z[i];
        z[j] = s / L[j + j * n];
    }
}

void orc_trmv_lower(const double *L, int64_t n, double *z) {
    for (int64_t i = n - 1; i >= 0; i--) {
        double s = 0.0;
        for (int64_t j = 0; j <= i; j++) s += L[i + j * n] * z[j];
        z[i] = s;
    }
}

double orc_dot(const double *a, const double *b, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

static int orc_variables(int kernel) { return kernel == 1 ? 2 : 1; }

int orc_loglik(int kernel, int64_t n_loc, int dts,
               const double *x, const double *y, const double *zc,
               const double *theta, int metric, double *z, double *C,
               double *loglik, double *logdet, double *quad,
               double *t_gen, double *t_facto, double *t_solve) {
    int64_t n = n_loc * orc_variables(kernel);
    double t0 = orc_now();
    /* The MLE always builds Sigma with metric 0 (ChameleonImplementation.cpp:168-173). */
    (void) metric;
    int st = orc_covariance_matrix(kernel, C, n, n, dts, 1, x, y, zc, x, y, zc, theta, 0);
    if (st) return st;
    double t1 = orc_now();
    if (orc_potrf(C, n, dts)) return 1;
    double t2 = orc_now();
    double ld = 2.0 * orc_logdet_half(C, n);
    orc_trsv_lower(C, n, z);
    double q = orc_dot(z, z, n);
    double t3 = orc_now();
    *logdet = ld;
    *quad = q;
    *loglik = -0.5 * q - 0.5 * ld - (double) (n / 2.0) * log(2.0 * ORC_PI);
    if (t_gen) *t_gen = t1 - t0;
    if (t_facto) *t_facto = t2 - t1;
    if (t_solve) *t_solve = t3 - t2;
    return 0;
}

int orc_generate_observations(int kernel, int64_t n_loc, int dts,
                              const double *x, const double *y, const double *zc,
                              const double *theta, int metric, int seed,
                              double *z_out, double *C) {
    int64_t n = n_loc * orc_variables(kernel);
    int iseed[4] = {seed, seed, seed, 1};
    orc_dlarnv_normal(iseed, n, z_out);
    int st = orc_covariance_matrix(kernel, C, n, n, dts, 1, x, y, zc, x, y, zc, theta, metric);
    if (st) return st;
    if (orc_potrf(C, n, dts)) return 1;
    orc_trmv_lower(C, n, z_out);
    return 0;
}

int orc_krige(int kernel, int64_t n_obs, int64_t n_miss, int dts,
              const double *xo, const double *yo, const double *xm, const double *ym,
              const double *theta, int metric, const double *zobs,
              const double *z_actual, double *zmiss, double *mspe) {
    (void) metric; /* prediction builds C22/C12 with metric 0 (LinearAlgebraMethods.cpp:530-537) */
    int p = orc_variables(kernel);
    int64_t no = n_obs * p, nm = n_miss * p;
    double *C22 = (double *) malloc(sizeof(double) * (size_t) (no * no));
    double *C12 = (double *) malloc(sizeof(double) * (size_t) (nm * no));
    double *x = (double *) malloc(sizeof(double) * (size_t) no);
    if (!C22 || !C12 || !x) {
        free(C22); free(C12); free(x);
        return -3;
    }
    int st = orc_covariance_matrix(kernel, C22, no, no, dts, 1, xo, yo, NULL, xo, yo, NULL, theta, 0);
    if (!st) st = orc_covariance_matrix(kernel, C12, nm, no, dts, 1, xm, ym, NULL, xo, yo, NULL, theta, 0);
    if (!st && orc_potrf(C22, no, dts)) st = 1;
    if (!st) {
        memcpy(x, zobs, sizeof(double) * (size_t) no);
        orc_trsv_lower(C22, no, x);
        orc_trsv_lower_trans(C22, no, x);
        for (int64_t i = 0; i < nm; i++) {
            double s = 0.0;
            for (int64_t j = 0; j < no; j++) s += C12[i + j * nm] * x[j];
            zmiss[i] = s;
        }
        if (z_actual) {
            double e = 0.0;
            for (int64_t i = 0; i < nm; i++) e += (z_actual[i] - zmiss[i]) * (z_actual[i] - zmiss[i]);
            *mspe = e / (double) nm;
        } else {
            *mspe = -1.0;
        }
    }
    free(C22);
    free(C12);
    free(x);
    return st;
}

void orc_mspe(int kernel, int64_t nm, int dts, const double *z_actual, const double *zmiss, double *out) {
    /* per-tile partials added in tile order (dmse-codelet.cpp:62-75;
     * dmse-bivariate-codelet.cpp:77-100: within a tile, even i -> error 1, odd i -> error 2) */
    double e = 0.0, e1 = 0.0, e2 = 0.0;
    for (int64_t t0 = 0; t0 < nm; t0 += dts) {
        int64_t rows = nm - t0 < dts ? nm - t0 : dts;
        double l = 0.0, l1 = 0.0, l2 = 0.0;
        for (int64_t i = 0; i < rows; i++) {
            double d = pow(z_actual[t0 + i] - zmiss[t0 + i], 2);
            if (i % 2 == 0) l1 += d;
            else l2 += d;
            l += d;
        }
        e += l;
        e1 += l1;
        e2 += l2;
    }
    /* LinearAlgebraMethods.cpp:575-577 */
    out[0] = e / (double) nm;
    if (orc_variables(kernel) == 2) {
        out[1] = e1 / (double) (nm / 2);
        out[2] = e2 / (double) (nm / 2);
    } else {
        out[1] = 0.0;
        out[2] = 0.0;
    }
}

void orc_shuffle(orc_rand_t *s, double *z, double *x, double *y, int n) {
    /* PredictionHelpers.cpp:141-168 */
    if (n <= 1) return;
    for (size_t i = 0; i < (size_t) (n - 1); i++) {
        size_t j = i + (size_t) (orc_rand(s) / (2147483647 / (n - (int) i) + 1));
        double t = z[j]; z[j] = z[i]; z[i] = t;
        t = x[j]; x[j] = x[i]; x[i] = t;
        t = y[j]; y[j] = y[i]; y[i] = t;
    }
}
