"""The library's BOBYQA (host code, no GPU) on analytic problems and on the
oracle's log-likelihood.

Pins: the optimiser's first evaluations (start, start + rho_beg e_1,
start + rho_beg e_2 with rho_beg = (ub - lb)/4) reproduce the reference's
modeling fixture -24.026000 (TestExaGeoStatApi.cpp:93, maxeval 3); the
converged estimate matches an independent optimiser's (scipy Nelder-Mead on the
same oracle objective, tests/golden/mle_n1600.json) within 1e-4.
"""
import json
import os

import numpy as np
import pytest

from exageostatcpp_amd import capi
from oracle import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_rosenbrock_converges():
    f = lambda v: 100 * (v[1] - v[0] ** 2) ** 2 + (1 - v[0]) ** 2
    x, fmin, nev, reason = capi.bobyqa(f, [-1.2, 1.0], [-2, -2], [2, 2], maxeval=3000, rho_end=1e-9)
    assert np.allclose(x, [1, 1], atol=1e-6) and fmin < 1e-12
    assert reason == 1 and nev < 400


def test_quadratic_with_active_bounds():
    c = np.array([-1.0, 2.0, 0.3, 0.7])
    f = lambda v: float(np.sum((v - c) ** 2) + 0.2 * v[2] * v[3])
    x, fmin, _, _ = capi.bobyqa(f, [0.5] * 4, [0] * 4, [1] * 4, maxeval=2000, rho_end=1e-9)
    # KKT point: x0 and x1 on their bounds, (x2, x3) solve the 2x2 system
    A = np.array([[2.0, 0.2], [0.2, 2.0]])
    x23 = np.linalg.solve(A, 2 * c[2:])
    assert np.allclose(x, [0, 1, x23[0], x23[1]], atol=1e-6)


def test_six_parameters_interior():
    w = np.arange(1, 7)
    t = 0.1 * w
    f = lambda v: float(np.sum(w * (v - t) ** 2) + 0.3 * v[0] * v[5])
    x, _, _, _ = capi.bobyqa(f, [-3] * 6, [-3] * 6, [3] * 6, maxeval=3000, rho_end=1e-9)
    H = np.diag(2.0 * w)
    H[0, 5] = H[5, 0] = 0.3
    assert np.allclose(x, np.linalg.solve(H, 2 * w * t), atol=1e-6)


def test_maxeval_and_ftol_stop():
    f = lambda v: float(np.sum((v - 0.3) ** 2))
    _, _, nev, reason = capi.bobyqa(f, [0, 0], [-1, -1], [1, 1], maxeval=4)
    assert nev == 4 and reason == 3
    quartic = lambda v: float(np.sum((v - 0.3) ** 4))
    _, _, nev2, reason2 = capi.bobyqa(quartic, [0, 0], [-1, -1], [1, 1], ftol_abs=1e-6, maxeval=500)
    assert reason2 == 2 and nev2 < 500


def test_invalid_bounds_and_nan():
    with pytest.raises(capi.EGSError):
        capi.bobyqa(lambda v: 0.0, [0.5], [1.0], [0.0])
    with pytest.raises(capi.EGSError):
        capi.bobyqa(lambda v: float("nan"), [0.5], [0.0], [1.0])


def test_first_evaluations_reproduce_modeling_fixture(fixtures):
    f = fixtures["modeling_exact"]
    L16 = fixtures["locations16"]
    loc = (np.array(L16["x"]), np.array(L16["y"]))
    z = np.array(L16["z"])
    seen = []

    def obj(v):
        seen.append(np.array(v))
        return -orc.loglik(0, loc, list(v), z, f["dts"])["llh"]

    lb, ub = np.array(f["lb"]), np.array(f["ub"])
    x, fmin, nev, reason = capi.bobyqa(obj, lb, lb, ub, ftol_abs=10.0 ** -f["tolerance"], maxeval=f["maxeval"])
    step = 0.25 * (ub - lb)
    assert nev == 3 and reason == 3
    assert np.allclose(seen[0], lb, atol=1e-14)
    assert np.allclose(seen[1], lb + step * [1, 0, 0], atol=1e-14)
    assert np.allclose(seen[2], lb + step * [0, 1, 0], atol=1e-14)
    assert abs(-fmin - f["loglik"]) < f["tol"]


def test_mle_n1600_on_the_oracle_matches_independent_optimum():
    with open(os.path.join(ROOT, "tests", "golden", "mle_n1600.json")) as fh:
        g = json.load(fh)
    x, y = orc.generate_locations(g["n"], seed=0)
    z = orc.generate_observations(0, (x, y), g["theta_true"], g["dts"], seed=0)
    obj = lambda v: -orc.loglik(0, (x, y), list(v), z, g["dts"])["llh"]
    th, fmin, nev, _ = capi.bobyqa(obj, g["lb"], g["lb"], g["ub"], ftol_abs=1e-9, maxeval=1000)
    assert np.max(np.abs(th - np.array(g["theta_hat"]))) < 1e-4
    assert -fmin >= g["loglik"] - 1e-6
