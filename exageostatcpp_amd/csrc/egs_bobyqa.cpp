// egs_bobyqa.cpp -- BOBYQA (Powell 2009) for the likelihood maximisation.
// Host code only.  The algorithm and its references are described in
// egs_bobyqa.hpp.
#include "egs_bobyqa.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace egs {
namespace {

using Vec = std::vector<double>;

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kPi = 3.14159265358979323846;

double dot(const double *a, const double *b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// Gauss-Jordan inversion with partial pivoting; false when a pivot vanishes.
bool invert(Vec &a, int N) {
    Vec inv(static_cast<size_t>(N) * N, 0.0);
    for (int i = 0; i < N; ++i) inv[i * N + i] = 1.0;
    double amax = 0.0;
    for (double v : a) amax = std::max(amax, std::fabs(v));
    if (amax == 0.0) return false;
    for (int c = 0; c < N; ++c) {
        int p = c;
        for (int r = c + 1; r < N; ++r)
            if (std::fabs(a[r * N + c]) > std::fabs(a[p * N + c])) p = r;
        if (std::fabs(a[p * N + c]) <= 1e-14 * amax) return false;
        if (p != c)
            for (int k = 0; k < N; ++k) {
                std::swap(a[p * N + k], a[c * N + k]);
                std::swap(inv[p * N + k], inv[c * N + k]);
            }
        const double d = 1.0 / a[c * N + c];
        for (int k = 0; k < N; ++k) {
            a[c * N + k] *= d;
            inv[c * N + k] *= d;
        }
        for (int r = 0; r < N; ++r) {
            if (r == c) continue;
            const double f = a[r * N + c];
            if (f == 0.0) continue;
            for (int k = 0; k < N; ++k) {
                a[r * N + k] -= f * a[c * N + k];
                inv[r * N + k] -= f * inv[c * N + k];
            }
        }
    }
    a.swap(inv);
    return true;
}

class Solver {
public:
    Solver(int n, const double *lb, const double *ub, const std::function<double(const double *)> &f,
           const BobyqaOptions &opt)
        : n_(n), m_(2 * n + 1), N_(3 * n + 2), f_(f), opt_(opt), scale_(n), lo_(n), hi_(n), lb_(lb, lb + n),
          ub_(ub, ub + n) {}

    BobyqaResult run(double *x);

private:
    // ---- objective in scaled coordinates --------------------------------------------------
    bool evaluate(const double *u, double &fval) {
        Vec xv(n_);
        for (int i = 0; i < n_; ++i) xv[i] = std::min(ub_[i], std::max(lb_[i], u[i] * scale_[i]));
        fval = f_(xv.data());
        ++nf_;
        if (fval < best_f_ || best_x_.empty()) {
            best_f_ = fval;
            best_x_ = xv;
        }
        return !std::isnan(fval);
    }
    bool budget_left() const { return nf_ < opt_.maxeval; }

    const double *pt(int k) const { return &y_[static_cast<size_t>(k) * n_]; }
    double *pt(int k) { return &y_[static_cast<size_t>(k) * n_]; }

    // ---- quadratic model centred at the best point ---------------------------------------
    double model_change(const double *s) const {  // Q(xopt + s) - Q(xopt)
        double v = dot(g_.data(), s, n_);
        for (int a = 0; a < n_; ++a) v += 0.5 * s[a] * dot(&H_[a * n_], s, n_);
        return v;
    }
    void recentre(const double *s) {  // move the centre to xopt + s
        q0_ += model_change(s);
        for (int a = 0; a < n_; ++a) g_[a] += dot(&H_[a * n_], s, n_);
    }

    // Inverse of the interpolation (KKT) matrix of the current points, in
    // coordinates (y - xopt)/sc.  Rows/columns: m Lagrange weights, constant,
    // n gradient terms.
    bool build_system() {
        const double *xo = pt(kopt_);
        sc_ = 0.0;
        for (int k = 0; k < m_; ++k) {
            double d2 = 0.0;
            for (int i = 0; i < n_; ++i) d2 += (pt(k)[i] - xo[i]) * (pt(k)[i] - xo[i]);
            sc_ = std::max(sc_, std::sqrt(d2));
        }
        if (sc_ == 0.0) return false;
        yp_.assign(static_cast<size_t>(m_) * n_, 0.0);
        for (int k = 0; k < m_; ++k)
            for (int i = 0; i < n_; ++i) yp_[k * n_ + i] = (pt(k)[i] - xo[i]) / sc_;
        W_.assign(static_cast<size_t>(N_) * N_, 0.0);
        for (int i = 0; i < m_; ++i) {
            for (int j = 0; j < m_; ++j) {
                const double p = dot(&yp_[i * n_], &yp_[j * n_], n_);
                W_[i * N_ + j] = 0.5 * p * p;
            }
            W_[i * N_ + m_] = W_[m_ * N_ + i] = 1.0;
            for (int t = 0; t < n_; ++t) W_[i * N_ + m_ + 1 + t] = W_[(m_ + 1 + t) * N_ + i] = yp_[i * n_ + t];
        }
        return invert(W_, N_);
    }

    // Quadratic (value at xopt, gradient, Hessian) with coefficient vector coef.
    void coeffs_to_quad(const Vec &coef, double &c, Vec &gq, Vec &Hq) const {
        c = coef[m_];
        gq.assign(n_, 0.0);
        Hq.assign(static_cast<size_t>(n_) * n_, 0.0);
        for (int t = 0; t < n_; ++t) gq[t] = coef[m_ + 1 + t] / sc_;
        const double inv2 = 1.0 / (sc_ * sc_);
        for (int j = 0; j < m_; ++j) {
            const double lam = coef[j] * inv2;
            if (lam == 0.0) continue;
            for (int a = 0; a < n_; ++a)
                for (int b = 0; b < n_; ++b) Hq[a * n_ + b] += lam * yp_[j * n_ + a] * yp_[j * n_ + b];
        }
    }

    // Least-Frobenius-norm correction so that the model interpolates every point.
    void refit() {
        const double *xo = pt(kopt_);
        Vec rhs(N_, 0.0), s(n_);
        for (int k = 0; k < m_; ++k) {
            for (int i = 0; i < n_; ++i) s[i] = pt(k)[i] - xo[i];
            rhs[k] = fv_[k] - (q0_ + model_change(s.data()));
        }
        Vec coef(N_, 0.0);
        for (int i = 0; i < N_; ++i) coef[i] = dot(&W_[i * N_], rhs.data(), N_);
        double c;
        Vec gq, Hq;
        coeffs_to_quad(coef, c, gq, Hq);
        q0_ += c;
        for (int a = 0; a < n_; ++a) g_[a] += gq[a];
        for (size_t i = 0; i < H_.size(); ++i) H_[i] += Hq[i];
    }

    // tau_k = l_k(xopt + s) for every k and beta, so that replacing point k by
    // xopt + s changes the interpolation determinant by sigma_k = alpha_k beta + tau_k^2.
    void denominators(const double *s, Vec &tau, double &beta) const {
        Vec xp(n_), w(N_, 0.0), Hw(N_, 0.0);
        for (int i = 0; i < n_; ++i) xp[i] = s[i] / sc_;
        for (int j = 0; j < m_; ++j) {
            const double p = dot(&yp_[j * n_], xp.data(), n_);
            w[j] = 0.5 * p * p;
        }
        w[m_] = 1.0;
        for (int t = 0; t < n_; ++t) w[m_ + 1 + t] = xp[t];
        for (int i = 0; i < N_; ++i) Hw[i] = dot(&W_[i * N_], w.data(), N_);
        const double xx = dot(xp.data(), xp.data(), n_);
        beta = 0.5 * xx * xx - dot(w.data(), Hw.data(), N_);
        tau.assign(Hw.begin(), Hw.begin() + m_);
    }
    double sigma(int k, const double *s) const {
        Vec tau;
        double beta;
        denominators(s, tau, beta);
        return W_[k * N_ + k] * beta + tau[k] * tau[k];
    }

    // ---- trust-region subproblem: min g's + s'Hs/2, |s| <= delta, bounds ----------------
    void trust_region_step(double delta, Vec &d, double &crvmin) const;
    // ---- model-improvement step for point knew within radius adelt ------------------------
    bool geometry_step(int knew, double adelt, Vec &s) const;

    // Replace point k by xopt + s with value fval; refit.  False if the point set degenerates.
    bool include(int k, const Vec &s, double fval) {
        const double *xo = pt(kopt_);
        Vec xnew(n_);
        for (int i = 0; i < n_; ++i) xnew[i] = std::min(hi_[i], std::max(lo_[i], xo[i] + s[i]));
        const bool better = fval < fv_[kopt_];
        if (better) {
            Vec sc(n_);
            for (int i = 0; i < n_; ++i) sc[i] = xnew[i] - xo[i];
            recentre(sc.data());
        }
        std::copy(xnew.begin(), xnew.end(), pt(k));
        fv_[k] = fval;
        if (better) kopt_ = k;
        if (!build_system()) return false;
        refit();
        return true;
    }

    // Fresh 2n+1 design around the best point with step `step` (used when the
    // interpolation set degenerates).  False when the budget ran out.
    bool rebuild(double step);

    int n_, m_, N_;
    const std::function<double(const double *)> &f_;
    BobyqaOptions opt_;
    Vec scale_, lo_, hi_, lb_, ub_;
    Vec y_, fv_, g_, H_, W_, yp_;
    double q0_ = 0.0, sc_ = 1.0;
    int kopt_ = 0, nf_ = 0;
    double best_f_ = kInf;
    Vec best_x_;
    int geo_knew_ = -1;  // point replaced by the pending model-improvement step
};

void Solver::trust_region_step(double delta, Vec &d, double &crvmin) const {
    const double *xo = pt(kopt_);
    d.assign(n_, 0.0);
    crvmin = -1.0;
    std::vector<char> fixed(n_, 0);
    for (int i = 0; i < n_; ++i)
        if ((xo[i] <= lo_[i] && g_[i] >= 0.0) || (xo[i] >= hi_[i] && g_[i] <= 0.0)) fixed[i] = 1;
    const double delsq = delta * delta;
    Vec gd(n_), r(n_), p(n_), Hp(n_);
    auto grad_at = [&](const Vec &dv, Vec &out) {
        for (int a = 0; a < n_; ++a) out[a] = g_[a] + dot(&H_[a * n_], dv.data(), n_);
    };
    bool boundary = false;
    const double g0 = dot(g_.data(), g_.data(), n_);
    // conjugate gradients on the free variables, restarted when a bound becomes active
    for (int restart = 0; restart <= n_ && !boundary; ++restart) {
        grad_at(d, gd);
        double rr = 0.0;
        for (int i = 0; i < n_; ++i) {
            r[i] = fixed[i] ? 0.0 : -gd[i];
            rr += r[i] * r[i];
        }
        if (rr <= 1e-30 * std::max(g0, 1e-300)) break;
        const double rr0 = rr;
        p = r;
        bool hit = false, done = false;
        for (int it = 0; it < n_; ++it) {
            for (int a = 0; a < n_; ++a) Hp[a] = dot(&H_[a * n_], p.data(), n_);
            const double pHp = dot(p.data(), Hp.data(), n_), pp = dot(p.data(), p.data(), n_);
            const double dp = dot(d.data(), p.data(), n_), rem = delsq - dot(d.data(), d.data(), n_);
            if (rem <= 0.0 || pp == 0.0) { boundary = rem <= 0.0; break; }
            const double t_tr = rem / (dp + std::sqrt(dp * dp + pp * rem));
            double t_bd = kInf;
            int ib = -1;
            for (int i = 0; i < n_; ++i) {
                if (fixed[i] || p[i] == 0.0) continue;
                double lim = p[i] > 0.0 ? (hi_[i] - xo[i] - d[i]) / p[i] : (lo_[i] - xo[i] - d[i]) / p[i];
                lim = std::max(lim, 0.0);
                if (lim < t_bd) { t_bd = lim; ib = i; }
            }
            const double t_cg = pHp > 0.0 ? rr / pHp : kInf;
            if (pHp > 0.0) crvmin = crvmin < 0.0 ? pHp / pp : std::min(crvmin, pHp / pp);
            const double t = std::min(t_cg, std::min(t_tr, t_bd));
            for (int i = 0; i < n_; ++i) {
                d[i] += t * p[i];
                r[i] -= t * Hp[i];
            }
            if (ib >= 0 && t == t_bd && t_bd < t_tr) {
                d[ib] = (p[ib] > 0.0 ? hi_[ib] : lo_[ib]) - xo[ib];
                fixed[ib] = 1;
                hit = true;
                break;
            }
            if (t == t_tr) { boundary = true; break; }
            double rr_new = 0.0;
            for (int i = 0; i < n_; ++i) {
                if (fixed[i]) r[i] = 0.0;
                rr_new += r[i] * r[i];
            }
            if (rr_new <= 1e-20 * rr0) { done = true; break; }
            for (int i = 0; i < n_; ++i) p[i] = fixed[i] ? 0.0 : r[i] + (rr_new / rr) * p[i];
            rr = rr_new;
        }
        if (done || !hit) break;
    }
    if (!boundary) return;
    crvmin = 0.0;
    // rotations d(theta) = d_fixed + cos(theta) d_free + sin(theta) t on the sphere |d| = delta
    auto qval = [&](const Vec &dv) {
        double v = dot(g_.data(), dv.data(), n_);
        for (int a = 0; a < n_; ++a) v += 0.5 * dv[a] * dot(&H_[a * n_], dv.data(), n_);
        return v;
    };
    const double q_start = 0.0;
    double q_cur = qval(d);
    Vec t(n_), dt(n_);
    for (int iter = 0; iter < 2 * n_ + 2; ++iter) {
        grad_at(d, gd);
        double ds = 0.0, dg = 0.0, gg = 0.0;
        for (int i = 0; i < n_; ++i) {
            if (fixed[i]) continue;
            ds += d[i] * d[i];
            dg += d[i] * gd[i];
            gg += gd[i] * gd[i];
        }
        if (ds <= 0.0) break;
        double temp = ds * gg - dg * dg;
        if (temp <= 1e-8 * ds * gg) break;
        temp = std::sqrt(temp);
        for (int i = 0; i < n_; ++i) t[i] = fixed[i] ? 0.0 : (dg * d[i] - ds * gd[i]) / temp;
        auto at = [&](double th, Vec &out) {
            const double c = std::cos(th), s = std::sin(th);
            for (int i = 0; i < n_; ++i) out[i] = fixed[i] ? d[i] : c * d[i] + s * t[i];
        };
        auto feasible = [&](const Vec &dv) {
            for (int i = 0; i < n_; ++i) {
                if (fixed[i]) continue;
                if (xo[i] + dv[i] > hi_[i] + 1e-15 * std::fabs(hi_[i]) ||
                    xo[i] + dv[i] < lo_[i] - 1e-15 * std::fabs(lo_[i]))
                    return false;
            }
            return true;
        };
        // largest feasible angle in (0, pi/2]
        const int kGrid = 96;
        double th_max = 0.5 * kPi;
        bool clipped = false;
        for (int s = 1; s <= kGrid; ++s) {
            const double th = 0.5 * kPi * s / kGrid;
            at(th, dt);
            if (!feasible(dt)) {
                double a = 0.5 * kPi * (s - 1) / kGrid, b = th;
                for (int it = 0; it < 50; ++it) {
                    const double mid = 0.5 * (a + b);
                    at(mid, dt);
                    (feasible(dt) ? a : b) = mid;
                }
                th_max = a;
                clipped = true;
                break;
            }
        }
        if (th_max <= 0.0) break;
        // minimise q over [0, th_max]: grid, then golden-section refinement
        double th_best = 0.0, q_best = q_cur;
        for (int s = 1; s <= 48; ++s) {
            const double th = th_max * s / 48;
            at(th, dt);
            const double q = qval(dt);
            if (q < q_best) { q_best = q; th_best = th; }
        }
        if (th_best > 0.0) {
            double a = std::max(0.0, th_best - th_max / 48), b = std::min(th_max, th_best + th_max / 48);
            const double gr = 0.5 * (std::sqrt(5.0) - 1.0);
            for (int it = 0; it < 40; ++it) {
                const double c1 = b - gr * (b - a), c2 = a + gr * (b - a);
                at(c1, dt);
                const double q1 = qval(dt);
                at(c2, dt);
                const double q2 = qval(dt);
                if (q1 < q2) b = c2; else a = c1;
            }
            const double th = 0.5 * (a + b);
            at(th, dt);
            const double q = qval(dt);
            if (q < q_best) { q_best = q; th_best = th; }
        }
        const double red = q_cur - q_best;
        if (th_best <= 0.0 || red <= 0.0) break;
        at(th_best, dt);
        d = dt;
        q_cur = q_best;
        if (clipped && th_best >= th_max * (1.0 - 1e-12)) {
            // the rotation stopped on a bound: fix the variable that reached it
            int ib = -1;
            double best = kInf;
            for (int i = 0; i < n_; ++i) {
                if (fixed[i]) continue;
                const double gap = std::min(std::fabs(xo[i] + d[i] - hi_[i]), std::fabs(xo[i] + d[i] - lo_[i]));
                if (gap < best) { best = gap; ib = i; }
            }
            if (ib >= 0) {
                d[ib] = std::fabs(xo[ib] + d[ib] - hi_[ib]) < std::fabs(xo[ib] + d[ib] - lo_[ib]) ? hi_[ib] - xo[ib]
                                                                                              : lo_[ib] - xo[ib];
                fixed[ib] = 1;
            }
        }
        if (red <= 0.01 * (q_start - q_cur)) break;
    }
}

bool Solver::geometry_step(int knew, double adelt, Vec &s) const {
    const double *xo = pt(kopt_);
    Vec e(N_, 0.0);
    for (int i = 0; i < N_; ++i) e[i] = W_[i * N_ + knew];
    double l0;
    Vec gl, Hl;
    coeffs_to_quad(e, l0, gl, Hl);
    std::vector<Vec> cand;
    Vec u(n_);
    // lines through xopt and the other interpolation points
    for (int k = 0; k < m_; ++k) {
        if (k == kopt_) continue;
        for (int i = 0; i < n_; ++i) u[i] = pt(k)[i] - xo[i];
        const double dist = std::sqrt(dot(u.data(), u.data(), n_));
        if (dist == 0.0) continue;
        double thi = adelt / dist, tlo = -thi;
        for (int i = 0; i < n_; ++i) {
            if (u[i] > 0.0) {
                thi = std::min(thi, (hi_[i] - xo[i]) / u[i]);
                tlo = std::max(tlo, (lo_[i] - xo[i]) / u[i]);
            } else if (u[i] < 0.0) {
                thi = std::min(thi, (lo_[i] - xo[i]) / u[i]);
                tlo = std::max(tlo, (hi_[i] - xo[i]) / u[i]);
            }
        }
        std::vector<double> ts = {tlo, thi};
        const double a1 = dot(gl.data(), u.data(), n_);
        double a2 = 0.0;
        for (int a = 0; a < n_; ++a) a2 += u[a] * dot(&Hl[a * n_], u.data(), n_);
        if (a2 != 0.0) {
            const double ts0 = -a1 / a2;
            if (ts0 > tlo && ts0 < thi) ts.push_back(ts0);
        }
        for (double tv : ts) {
            if (tv == 0.0) continue;
            Vec c(n_);
            for (int i = 0; i < n_; ++i) c[i] = tv * u[i];
            cand.push_back(c);
        }
    }
    // bounded Cauchy steps along +-grad l_knew
    for (int sgn = -1; sgn <= 1; sgn += 2) {
        Vec c(n_, 0.0);
        std::vector<char> fixed(n_, 0);
        for (int it = 0; it <= n_; ++it) {
            double gg = 0.0, cc = 0.0;
            for (int i = 0; i < n_; ++i) {
                if (fixed[i]) cc += c[i] * c[i];
                else gg += gl[i] * gl[i];
            }
            const double rem = adelt * adelt - cc;
            if (gg == 0.0 || rem <= 0.0) break;
            const double len = std::sqrt(rem / gg);
            bool clip = false;
            for (int i = 0; i < n_; ++i) {
                if (fixed[i]) continue;
                const double v = sgn * len * gl[i];
                if (xo[i] + v > hi_[i]) { c[i] = hi_[i] - xo[i]; fixed[i] = 1; clip = true; }
                else if (xo[i] + v < lo_[i]) { c[i] = lo_[i] - xo[i]; fixed[i] = 1; clip = true; }
            }
            if (!clip) {
                for (int i = 0; i < n_; ++i)
                    if (!fixed[i]) c[i] = sgn * len * gl[i];
                break;
            }
        }
        if (dot(c.data(), c.data(), n_) > 0.0) cand.push_back(c);
    }
    double best = 0.0;
    for (const Vec &c : cand) {
        const double sg = std::fabs(sigma(knew, c.data()));
        if (sg > best) { best = sg; s = c; }
    }
    return best > 1e-12;
}

bool Solver::rebuild(double step) {
    Vec xo(pt(kopt_), pt(kopt_) + n_);
    const double fo = fv_[kopt_];
    Vec ny(static_cast<size_t>(m_) * n_), nfv(m_);
    std::copy(xo.begin(), xo.end(), ny.begin());
    nfv[0] = fo;
    for (int i = 0; i < n_; ++i) {
        double sa = step, sb = -step;
        if (xo[i] + sa > hi_[i]) { sa = -step; sb = xo[i] - 2 * step >= lo_[i] ? -2 * step : lo_[i] - xo[i]; }
        else if (xo[i] + sb < lo_[i]) { sb = xo[i] + 2 * step <= hi_[i] ? 2 * step : hi_[i] - xo[i]; }
        for (int w = 0; w < 2; ++w) {
            const int k = 1 + i + w * n_;
            std::copy(xo.begin(), xo.end(), ny.begin() + static_cast<size_t>(k) * n_);
            ny[static_cast<size_t>(k) * n_ + i] += w == 0 ? sa : sb;
        }
    }
    for (int k = 1; k < m_; ++k) {
        if (!budget_left()) return false;
        if (!evaluate(&ny[static_cast<size_t>(k) * n_], nfv[k])) return false;
    }
    y_.swap(ny);
    fv_.swap(nfv);
    kopt_ = static_cast<int>(std::min_element(fv_.begin(), fv_.end()) - fv_.begin());
    q0_ = fv_[kopt_];
    std::fill(g_.begin(), g_.end(), 0.0);
    std::fill(H_.begin(), H_.end(), 0.0);
    if (!build_system()) return false;
    refit();
    return true;
}

BobyqaResult Solver::run(double *x) {
    // initial steps and scaling: every variable moves by rho_beg = 1 in scaled units
    for (int i = 0; i < n_; ++i) {
        if (!(ub_[i] > lb_[i]) || !std::isfinite(lb_[i]) || !std::isfinite(ub_[i]))
            return {BobyqaStatus::kInvalid, kInf, 0};
        const double xi = std::min(ub_[i], std::max(lb_[i], x[i]));
        double step = opt_.step_scale * (ub_[i] - lb_[i]);
        if (ub_[i] - xi < step && ub_[i] > xi) step = 0.75 * (ub_[i] - xi);
        if (xi - lb_[i] < step && xi > lb_[i]) step = 0.75 * (xi - lb_[i]);
        scale_[i] = step;
        lo_[i] = lb_[i] / step;
        hi_[i] = ub_[i] / step;
        x[i] = xi;
    }
    const double rhobeg = 1.0;
    // start point moved onto a bound, or rho_beg away from it (Powell's initialisation)
    Vec x0(n_), sl(n_), su(n_);
    for (int i = 0; i < n_; ++i) {
        x0[i] = x[i] / scale_[i];
        sl[i] = lo_[i] - x0[i];
        su[i] = hi_[i] - x0[i];
        if (sl[i] >= -rhobeg) {
            if (sl[i] >= 0.0) { x0[i] = lo_[i]; sl[i] = 0.0; su[i] = hi_[i] - lo_[i]; }
            else { x0[i] = lo_[i] + rhobeg; sl[i] = -rhobeg; su[i] = std::max(hi_[i] - x0[i], rhobeg); }
        } else if (su[i] <= rhobeg) {
            if (su[i] <= 0.0) { x0[i] = hi_[i]; sl[i] = lo_[i] - hi_[i]; su[i] = 0.0; }
            else { x0[i] = hi_[i] - rhobeg; sl[i] = std::min(lo_[i] - x0[i], -rhobeg); su[i] = rhobeg; }
        }
    }
    y_.assign(static_cast<size_t>(m_) * n_, 0.0);
    fv_.assign(m_, kInf);
    g_.assign(n_, 0.0);
    H_.assign(static_cast<size_t>(n_) * n_, 0.0);
    for (int k = 0; k < m_; ++k) std::copy(x0.begin(), x0.end(), pt(k));
    for (int i = 0; i < n_; ++i) {
        pt(1 + i)[i] += su[i] == 0.0 ? -rhobeg : rhobeg;
        double sb = -rhobeg;
        if (sl[i] == 0.0) sb = std::min(2.0 * rhobeg, su[i]);
        if (su[i] == 0.0) sb = std::max(-2.0 * rhobeg, sl[i]);
        pt(1 + n_ + i)[i] += sb;
    }
    auto finish = [&](BobyqaStatus st) {
        std::copy(best_x_.begin(), best_x_.end(), x);
        return BobyqaResult{st, best_f_, nf_};
    };
    for (int k = 0; k < m_; ++k) {
        if (!evaluate(pt(k), fv_[k])) return finish(BobyqaStatus::kFailure);
        if (!budget_left()) return finish(BobyqaStatus::kMaxEval);
    }
    kopt_ = static_cast<int>(std::min_element(fv_.begin(), fv_.end()) - fv_.begin());
    q0_ = fv_[kopt_];
    if (!build_system()) return finish(BobyqaStatus::kFailure);
    refit();

    const double rhoend = std::min(opt_.rho_end, rhobeg);
    double rho = rhobeg, delta = rhobeg, ratio = 1.0, dnorm = 0.0;
    double diffa = 0.0, diffb = 0.0, diffc = 0.0;
    int ntrits = 0, nfsav = nf_;
    enum class Next { kTrust, kGeometry, kRho } next = Next::kTrust;
    double distsq = 0.0;
    Vec d(n_), s(n_), pending;
    double crvmin = 0.0;

    // evaluate xopt + step, include it as point knew (or let choose() pick), book-keep the prediction error
    auto try_point = [&](const Vec &step, bool trust, BobyqaStatus &stop) -> int {
        // returns: 0 continue, 1 stop
        const double fopt = fv_[kopt_];
        const double vquad = model_change(step.data());
        Vec xn(n_);
        for (int i = 0; i < n_; ++i) xn[i] = std::min(hi_[i], std::max(lo_[i], pt(kopt_)[i] + step[i]));
        double fnew;
        if (!evaluate(xn.data(), fnew)) { stop = BobyqaStatus::kFailure; return 1; }
        const double diff = fnew - fopt - vquad;
        diffc = diffb;
        diffb = diffa;
        diffa = std::fabs(diff);
        if (dnorm > rho) nfsav = nf_;
        int knew = -1;
        if (trust) {
            ratio = vquad < 0.0 ? (fnew - fopt) / vquad : -1.0;
            if (ratio <= 0.1) delta = std::min(0.5 * delta, dnorm);
            else if (ratio <= 0.7) delta = std::max(0.5 * delta, dnorm);
            else delta = std::max(0.5 * delta, 2.0 * dnorm);
            if (delta <= 1.5 * rho) delta = rho;
            Vec tau;
            double beta;
            denominators(step.data(), tau, beta);
            const double delsq = delta * delta;
            double scaden = 0.0;
            for (int k = 0; k < m_; ++k) {
                if (k == kopt_ && !(fnew < fopt)) continue;
                const double den = W_[k * N_ + k] * beta + tau[k] * tau[k];
                double d2 = 0.0;
                for (int i = 0; i < n_; ++i) d2 += (pt(k)[i] - pt(kopt_)[i]) * (pt(k)[i] - pt(kopt_)[i]);
                const double temp = std::max(1.0, (d2 / delsq) * (d2 / delsq));
                if (temp * den > scaden) { scaden = temp * den; knew = k; }
            }
            if (knew >= 0 && scaden <= 1e-12) knew = -1;
        } else {
            knew = geo_knew_;
        }
        if (knew < 0) {
            // no safe replacement: keep the set, or rebuild it around a better point
            if (fnew < fopt) {
                std::copy(xn.begin(), xn.end(), pt(kopt_));
                fv_[kopt_] = fnew;
                if (!budget_left()) { stop = BobyqaStatus::kMaxEval; return 1; }
                if (!rebuild(rho)) { stop = budget_left() ? BobyqaStatus::kFailure : BobyqaStatus::kMaxEval; return 1; }
            }
        } else {
            Vec st(n_);
            for (int i = 0; i < n_; ++i) st[i] = xn[i] - pt(kopt_)[i];
            if (!include(knew, st, fnew)) {
                if (!budget_left()) { stop = BobyqaStatus::kMaxEval; return 1; }
                if (!rebuild(rho)) { stop = budget_left() ? BobyqaStatus::kFailure : BobyqaStatus::kMaxEval; return 1; }
            }
        }
        if (fnew < fopt && opt_.ftol_abs > 0.0 && fopt - fnew < opt_.ftol_abs) { stop = BobyqaStatus::kFtol; return 1; }
        if (!budget_left()) { stop = BobyqaStatus::kMaxEval; return 1; }
        return 0;
    };

    BobyqaStatus stop = BobyqaStatus::kRhoEnd;
    for (;;) {
        if (next == Next::kTrust) {
            trust_region_step(delta, d, crvmin);
            dnorm = std::min(delta, std::sqrt(dot(d.data(), d.data(), n_)));
            if (dnorm < 0.5 * rho) {
                ntrits = -1;
                pending = d;
                distsq = (10.0 * rho) * (10.0 * rho);
                if (nf_ <= nfsav + 2) { next = Next::kGeometry; continue; }
                const double errbig = std::max(diffa, std::max(diffb, diffc));
                const double frhosq = 0.125 * rho * rho;
                if (crvmin > 0.0 && errbig > frhosq * crvmin) { next = Next::kGeometry; continue; }
                const double bdtol = errbig / rho;
                bool geo = false;
                for (int j = 0; j < n_; ++j) {
                    const double xj = pt(kopt_)[j] + d[j];
                    const double gj = g_[j] + dot(&H_[j * n_], d.data(), n_);
                    double bdtest = bdtol;
                    if (xj <= lo_[j]) bdtest = gj;
                    if (xj >= hi_[j]) bdtest = -gj;
                    if (bdtest < bdtol) {
                        bdtest += 0.5 * H_[j * n_ + j] * rho;
                        if (bdtest < bdtol) geo = true;
                    }
                }
                next = geo ? Next::kGeometry : Next::kRho;
                continue;
            }
            ++ntrits;
            pending.clear();
            if (try_point(d, true, stop)) break;
            if (ratio >= 0.1) { next = Next::kTrust; continue; }
            distsq = std::max(4.0 * delta * delta, 100.0 * rho * rho);
            next = Next::kGeometry;
            continue;
        }
        if (next == Next::kGeometry) {
            int knew = -1;
            double far = distsq;
            for (int k = 0; k < m_; ++k) {
                double d2 = 0.0;
                for (int i = 0; i < n_; ++i) d2 += (pt(k)[i] - pt(kopt_)[i]) * (pt(k)[i] - pt(kopt_)[i]);
                if (d2 > far) { far = d2; knew = k; }
            }
            if (knew >= 0) {
                const double dist = std::sqrt(far);
                if (ntrits == -1) {
                    delta = std::min(0.1 * delta, 0.5 * dist);
                    if (delta <= 1.5 * rho) delta = rho;
                }
                ntrits = 0;
                pending.clear();
                const double adelt = std::max(std::min(0.1 * dist, delta), rho);
                if (!geometry_step(knew, adelt, s)) {
                    if (!rebuild(rho)) { stop = budget_left() ? BobyqaStatus::kFailure : BobyqaStatus::kMaxEval; break; }
                    next = Next::kTrust;
                    continue;
                }
                dnorm = std::sqrt(dot(s.data(), s.data(), n_));
                geo_knew_ = knew;
                if (try_point(s, false, stop)) break;
                next = Next::kTrust;
                continue;
            }
            if (ntrits == -1) { next = Next::kRho; continue; }
            if (ratio > 0.0 || std::max(delta, dnorm) > rho) { next = Next::kTrust; continue; }
            next = Next::kRho;
            continue;
        }
        // next == kRho
        if (rho > rhoend) {
            delta = 0.5 * rho;
            const double r = rho / rhoend;
            if (r <= 16.0) rho = rhoend;
            else if (r <= 250.0) rho = std::sqrt(r) * rhoend;
            else rho = 0.1 * rho;
            delta = std::max(delta, rho);
            ntrits = 0;
            nfsav = nf_;
            next = Next::kTrust;
            continue;
        }
        // final: try the last short step once if it was never evaluated
        if (ntrits == -1 && !pending.empty() && dot(pending.data(), pending.data(), n_) > 0.0 && budget_left()) {
            Vec xn(n_);
            for (int i = 0; i < n_; ++i) xn[i] = std::min(hi_[i], std::max(lo_[i], pt(kopt_)[i] + pending[i]));
            double fnew;
            if (!evaluate(xn.data(), fnew)) { stop = BobyqaStatus::kFailure; break; }
        }
        stop = BobyqaStatus::kRhoEnd;
        break;
    }
    return finish(stop);
}

}  // namespace

BobyqaResult bobyqa_minimize(int n, double *x, const double *lb, const double *ub,
                             const std::function<double(const double *)> &f, const BobyqaOptions &opt) {
    if (n < 1 || n > 64 || opt.maxeval < 1) return {BobyqaStatus::kInvalid, kInf, 0};
    Solver s(n, lb, ub, f, opt);
    return s.run(x);
}

}  // namespace egs
