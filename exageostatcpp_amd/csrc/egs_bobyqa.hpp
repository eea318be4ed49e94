// egs_bobyqa.hpp -- bound-constrained derivative-free minimisation (BOBYQA).
//
// The reference maximises the log-likelihood with NLopt 2.7.1's LN_BOBYQA
// (src/runtime-solver/concrete/StarpuRuntimeSolver.cpp:38-48): lower/upper
// bounds, ftol_abs = 10^-tolerance, maxeval = max_mle_iterations, start at the
// lower bounds (src/configurations/Configurations.cpp:200).  NLopt is not part
// of the reference tree; this is an implementation of Powell's published
// algorithm (M.J.D. Powell, "The BOBYQA algorithm for bound constrained
// optimization without derivatives", Cambridge NA report NA2009/06):
//
//  * 2n+1 interpolation points, the first 2n+1 placed as in Powell's
//    initialisation (x0, x0 + rho_beg e_i, then x0 - rho_beg e_i, or
//    x0 + 2 rho_beg e_i when x0 sits on the lower bound), so the first
//    evaluations are x0, x0 + rho_beg e_1, x0 + rho_beg e_2, ...;
//  * a quadratic model updated by the least Frobenius norm change of its
//    Hessian that interpolates every point;
//  * trust-region steps from a truncated conjugate-gradient solver with an
//    active set for the bounds and rotations along the trust-region boundary;
//  * model-improvement steps that maximise the updating denominator
//    sigma = alpha*beta + tau^2 of the replaced point over lines through the
//    interpolation points and a bounded Cauchy step;
//  * Powell's rho schedule (rho <- rho/10, sqrt(rho*rho_end) near the end).
//
// The interpolation system (at most 20 x 20 for six parameters) is solved
// directly at every update instead of carrying Powell's factored inverse: each
// objective evaluation here is a dense factorisation on the GPU, so the
// O(m^3) host work per step is negligible, and the refit cannot drift.
//
// Initial step: NLopt's default for bounded variables, 0.25 (ub - lb), cut to
// 0.75 of the distance to a bound when the start lies strictly inside and
// closer than that.  Variables are scaled so every initial step is rho_beg.
#pragma once

#include <functional>
#include <vector>

namespace egs {

enum class BobyqaStatus : int {
    kRhoEnd = 1,     // trust-region radius reached rho_end
    kFtol = 2,       // an improvement of the best value was below ftol_abs
    kMaxEval = 3,    // evaluation budget spent
    kInvalid = -2,   // bounds or sizes invalid
    kFailure = -1,   // objective returned NaN
};

struct BobyqaOptions {
    double ftol_abs = 0.0;   // stop when an improvement of the best value is below this (0: off)
    int maxeval = 1000;      // objective evaluations, counting the initial design
    double rho_end = 1e-7;   // final trust-region radius, relative to the initial step
    double step_scale = 0.25;   // initial step as a fraction of ub - lb (NLopt's default 0.25)
};

struct BobyqaResult {
    BobyqaStatus status;
    double fmin;
    int nevals;
};

// Minimise f over lb <= x <= ub starting from x (updated in place to the best
// point).  f receives a point inside the bounds.
BobyqaResult bobyqa_minimize(int n, double *x, const double *lb, const double *ub,
                             const std::function<double(const double *)> &f, const BobyqaOptions &opt);

}  // namespace egs
