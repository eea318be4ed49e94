// FP64 Matern covariance element on the device.
//
// K_nu(x) (the reference calls GSL's gsl_sf_bessel_Knu at
// UnivariateMaternStationary.cpp:61-62 and BivariateMaternParsimonious.cpp:91-101):
//   nu = N + mu with |mu| <= 1/2;  K_mu and K_{mu+1} from
//     x <= 2 : N. M. Temme's power series (J. Comput. Phys. 19, 1975), with
//              Gamma_1(mu), Gamma_2(mu) summed from the Taylor coefficients of
//              1/Gamma(1+t) (Abramowitz & Stegun 6.1.34, evaluated to 25 digits);
//     x >  2 : Steed's evaluation of the continued fraction CF2 and its
//              normalising sum (I. J. Thompson & A. R. Barnett, J. Comput. Phys.
//              64, 1986/87);
//   then K_{v+1} = K_{v-1} + (2v/x) K_v upward to nu.
// Half-integer orders 1/2, 3/2, 5/2 use their closed forms.
#pragma once

#include <hip/hip_runtime.h>

namespace egs {

__device__ __forceinline__ double rgamma1p_even(double m2) {
    // sum_{k even} a_k m^k, a_k the Taylor coefficients of 1/Gamma(1+t)
    double s = 2.73603004860799984483151e-23;
    s = s * m2 + 1.337351730493693114864781e-22;
    s = s * m2 + -2.298745684435370206592479e-19;
    s = s * m2 + 1.186692254751600332579777e-18;
    s = s * m2 + 1.226778628238260790158894e-15;
    s = s * m2 + -2.05832605356650678322243e-14;
    s = s * m2 + -3.696805618642205708187816e-12;
    s = s * m2 + 1.04342671169110051049154e-10;
    s = s * m2 + 5.002007644469222930055665e-9;
    s = s * m2 + -2.056338416977607103450154e-7;
    s = s * m2 + -1.250493482142670657345359e-6;
    s = s * m2 + 1.280502823881161861531986e-4;
    s = s * m2 + -1.165167591859065112113971e-3;
    s = s * m2 + -9.621971527876973562114922e-3;
    s = s * m2 + 0.1665386113822914895017008;
    s = s * m2 + -0.6558780715202538810770195;
    s = s * m2 + 1.0;
    return s;
}

__device__ __forceinline__ double rgamma1p_odd(double m2) {
    // sum_{k odd} a_k m^(k-1)
    double s = -2.054233551766672789325025e-22;
    s = s * m2 + 1.714406321927337433383963e-20;
    s = s * m2 + 1.412380655318031781555804e-18;
    s = s * m2 + -1.181259301697458769513765e-16;
    s = s * m2 + -5.348122539423017982370017e-15;
    s = s * m2 + 5.100370287454475979015481e-13;
    s = s * m2 + 7.782263439905071254049937e-12;
    s = s * m2 + -1.181274570487020144588127e-9;
    s = s * m2 + 6.116095104481415817862499e-9;
    s = s * m2 + 1.13302723198169588237413e-6;
    s = s * m2 + -2.013485478078823865568939e-5;
    s = s * m2 + -2.1524167411495097281573e-4;
    s = s * m2 + 7.21894324666309954239501e-3;
    s = s * m2 + -0.0421977345555443367482083;
    s = s * m2 + -0.04200263503409523552900393;
    s = s * m2 + 0.5772156649015328606065121;
    return s;
}

// K_mu(x), K_{mu+1}(x) for |mu| <= 1/2, 0 < x <= 2.
__device__ inline void bessel_k_series(double mu, double x, double &k_mu, double &k_mu1) {
    const double pi = 3.141592653589793238462643;
    const double hx = 0.5 * x;
    const double ln2x = -log(hx);
    const double sg = mu * ln2x;
    const double m2 = mu * mu;
    const double g1 = -rgamma1p_odd(m2);
    const double g2 = rgamma1p_even(m2);
    const double inv_gp = g2 - mu * g1;   // 1 / Gamma(1 + mu)
    const double inv_gm = g2 + mu * g1;   // 1 / Gamma(1 - mu)
    const double pim = pi * mu;
    const double fac = (mu == 0.0) ? 1.0 : pim / sin(pim);
    const double shs = (sg == 0.0) ? 1.0 : sinh(sg) / sg;
    const double es = exp(sg);
    double f = fac * (g1 * cosh(sg) + g2 * shs * ln2x);
    double p = 0.5 * es / inv_gp;
    double q = 0.5 / (es * inv_gm);
    double c = 1.0;
    const double hx2 = hx * hx;
    double s0 = f, s1 = p;
    for (int k = 1; k < 200; ++k) {
        const double dk = (double) k;
        f = (dk * f + p + q) / (dk * dk - m2);
        c *= hx2 / dk;
        p /= (dk - mu);
        q /= (dk + mu);
        const double t0 = c * f;
        s0 += t0;
        s1 += c * (p - dk * f);
        if (fabs(t0) < 1e-17 * fabs(s0)) break;
    }
    k_mu = s0;
    k_mu1 = s1 * (2.0 / x);
}

// K_mu(x), K_{mu+1}(x) for |mu| <= 1/2, x > 2.
__device__ inline void bessel_k_cf2(double mu, double x, double &k_mu, double &k_mu1) {
    const double pi = 3.141592653589793238462643;
    const double a1 = 0.25 - mu * mu;
    double b = 2.0 * (1.0 + x);
    double d = 1.0 / b;
    double h = d, dh = d;
    double q1 = 0.0, q2 = 1.0;
    double q = a1, c = a1, a = -a1;
    double s = 1.0 + q * dh;
    for (int i = 2; i < 2000; ++i) {
        a -= 2.0 * (i - 1);
        c = -a * c / i;
        const double qn = (q1 - b * q2) / a;
        q1 = q2;
        q2 = qn;
        q += c * qn;
        b += 2.0;
        d = 1.0 / (b + a * d);
        dh = (b * d - 1.0) * dh;
        h += dh;
        const double ds = q * dh;
        s += ds;
        if (fabs(ds) < 1e-17 * fabs(s)) break;
    }
    h *= a1;
    const double k0 = sqrt(pi / (2.0 * x)) * exp(-x) / s;
    k_mu = k0;
    k_mu1 = k0 * (mu + x + 0.5 - h) / x;
}

__device__ inline double bessel_knu(double nu, double x) {
    // orders are validated on the host (0 < nu <= 100); anything else is NaN
    if (!(nu >= 0.0 && nu <= 100.0) || !(x > 0.0)) return __builtin_nan("");
    const int nl = (int) floor(nu + 0.5);
    const double mu = nu - nl;
    double k0, k1;
    if (x <= 2.0) bessel_k_series(mu, x, k0, k1);
    else bessel_k_cf2(mu, x, k0, k1);
    for (int i = 1; i <= nl; ++i) {
        const double k2 = (mu + i) * (2.0 / x) * k1 + k0;
        k0 = k1;
        k1 = k2;
    }
    return k0;
}

// Matern correlation scaled by con = sigma^2 / (2^(nu-1) Gamma(nu)), evaluated at
// the scaled distance d = dist / range > 0:  con * d^nu * K_nu(d).
__device__ inline double matern_scaled(double d, double con, double sigma2, double nu) {
    if (nu == 0.5) return sigma2 * exp(-d);
    if (nu == 1.5) return sigma2 * (1.0 + d) * exp(-d);
    if (nu == 2.5) return sigma2 * (1.0 + d + d * d * (1.0 / 3.0)) * exp(-d);
    return con * pow(d, nu) * bessel_knu(nu, d);
}

// Great-circle distance in km with x = latitude, y = longitude in degrees
// (DistanceCalculationHelpers.cpp:27-36; earth radius 6371 km).
__device__ inline double distance_earth(double lat1, double lon1, double lat2, double lon2) {
    const double deg = 3.141592653589793 / 180.0;
    const double l1 = lat1 * deg, l2 = lat2 * deg;
    const double u = sin((l2 - l1) * 0.5);
    const double v = sin((lon2 * deg - lon1 * deg) * 0.5);
    return 2.0 * 6371.0 * asin(sqrt(u * u + cos(l1) * cos(l2) * v * v));
}

// Locations as structure of arrays; z == nullptr for 2D.
struct Locs {
    const double *x;
    const double *y;
    const double *z;
};

__device__ __forceinline__ double distance(double x1, double y1, double z1, double x2, double y2, double z2,
                                           int metric, bool three_d) {
    if (metric == 1) return distance_earth(x1, y1, x2, y2);
    const double dx = x2 - x1, dy = y2 - y1;
    if (!three_d) return sqrt(dx * dx + dy * dy);
    const double dz = z2 - z1;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Kernel parameters prepared on the host from theta (the gamma-function
// constants are evaluated once there).
//   kernel 0 (UnivariateMaternStationary, UnivariateMaternStationary.cpp:47-65):
//     sigma2 = theta0, range = theta1, nu = theta2, con = sigma2/(2^(nu-1) Gamma(nu)).
//   kernel 1 (BivariateMaternParsimonious, BivariateMaternParsimonious.cpp:50-107):
//     per variable pair (a, b) in {0,1}^2: sig[a][b], nu[a][b], con[a][b], with the
//     cross terms using nu12 = (nu1 + nu2)/2 and rho * sqrt(s1 s2).
struct MaternParams {
    int kernel;
    int metric;
    double range;
    double sig[4];   // [a*2+b]
    double nu[4];
    double con[4];
};

// Covariance between matrix rows r and c (indices into the n-dimensional
// matrix; for the bivariate kernel, row r is variable r%2 at location r/2).
__device__ __forceinline__ double covariance_entry(const MaternParams &p, const Locs &l1, const Locs &l2,
                                                   int64_t r, int64_t c, bool three_d) {
    int64_t i = r, j = c;
    int pair = 0;
    if (p.kernel == 1) {
        i = r >> 1;
        j = c >> 1;
        pair = (int) ((r & 1) * 2 + (c & 1));
    }
    const double dist = distance(l1.x[i], l1.y[i], three_d ? l1.z[i] : 0.0, l2.x[j], l2.y[j],
                                 three_d ? l2.z[j] : 0.0, p.metric, three_d);
    const double d = dist / p.range;
    if (d == 0.0) return p.sig[pair];
    return matern_scaled(d, p.con[pair], p.sig[pair], p.nu[pair]);
}

}  // namespace egs
