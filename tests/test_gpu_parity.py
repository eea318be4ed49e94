"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Sizes keep the oracle within seconds; tolerances are written per test.  The
reference's fixtures (tests/golden/reference_fixtures.json) are checked on the
GPU path as well.
"""
import sys

import numpy as np
import pytest

import exageostatcpp_amd as egs
from oracle import orc

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = egs.Context(0)
    yield c
    c.close()


def _rng_locs(n, seed):
    rng = np.random.default_rng(seed)
    return rng.random(n), rng.random(n)


# ------------------------------------------------------------------ covariance
@pytest.mark.parametrize("theta", [[1.0, 0.1, 0.5], [0.7, 0.05, 1.0], [1.3, 0.2, 1.234], [2.0, 0.08, 2.5],
                                   [1.0, 0.3, 0.3], [1.1, 0.02, 4.2]])
def test_covariance_tile_univariate(ctx, theta):
    log("covariance tile univariate theta=%s" % theta)
    x, y = _rng_locs(300, 1)
    G = ctx.covariance_tile(0, 70, 90, 13, 201, (x, y), (x, y), theta)
    O = orc.covariance_tile(0, 70, 90, 13, 201, (x, y), (x, y), theta)
    assert np.max(np.abs(G - O) / np.abs(O)) < 1e-13


def test_covariance_tile_bivariate_3d_gcd(ctx):
    log("covariance tile bivariate / 3D / great-circle")
    x, y = _rng_locs(100, 2)
    th = [1.0, 1.5, 0.1, 0.5, 1.0, 0.5]
    G = ctx.covariance_tile(1, 40, 60, 20, 100, (x, y), (x, y), th)
    O = orc.covariance_tile(1, 40, 60, 20, 100, (x, y), (x, y), th)
    assert np.max(np.abs(G - O)) < 1e-13
    z = np.random.default_rng(3).random(100)
    G = ctx.covariance_tile(0, 50, 50, 0, 50, (x, y, z), (x, y, z), [1.0, 0.1, 1.7])
    O = orc.covariance_tile(0, 50, 50, 0, 50, (x, y, z), (x, y, z), [1.0, 0.1, 1.7])
    assert np.max(np.abs(G - O) / np.abs(O)) < 1e-13
    lat, lon = 180 * x - 90, 360 * y - 180
    G = ctx.covariance_tile(0, 60, 60, 0, 40, (lat, lon), (lat, lon), [1.0, 637.1, 0.5], metric=1)
    O = orc.covariance_tile(0, 60, 60, 0, 40, (lat, lon), (lat, lon), [1.0, 637.1, 0.5], metric=1)
    assert np.max(np.abs(G - O) / np.abs(O)) < 1e-12
    with pytest.raises(egs.EGSError, match="Great Circle"):
        ctx.covariance_tile(0, 5, 5, 0, 0, (x, y, z), (x, y, z), [1.0, 0.1, 0.5], metric=1)


@pytest.mark.parametrize("n,dts,kernel", [(300, 64, 0), (515, 100, 0), (130, 64, 1)])
def test_problem_covariance_matches_oracle(ctx, n, dts, kernel):
    log("problem covariance n=%d dts=%d kernel=%d" % (n, dts, kernel))
    x, y = _rng_locs(n, n)
    th = [1.0, 0.1, 0.9] if kernel == 0 else [1.0, 0.8, 0.1, 0.5, 1.2, 0.3]
    p = egs.Problem(ctx, kernel, x, y, dts=dts)
    p.covariance(th)
    G = p.get_matrix()
    O = np.tril(orc.covariance_matrix(kernel, (x, y), (x, y), th, dts=dts, lower=False))
    assert np.max(np.abs(G - O)) < 1e-14 * max(th[0], th[1])


# ------------------------------------------------------------------ factorization
@pytest.mark.parametrize("n,dts", [(64, 64), (200, 64), (700, 128), (1000, 320), (2100, 960)])
def test_potrf_matches_oracle(ctx, n, dts):
    log("potrf n=%d dts=%d" % (n, dts))
    x, y = orc.generate_locations(n, seed=5)
    th = [1.0, 0.1, 0.5]
    p = egs.Problem(ctx, 0, x, y, dts=dts)
    p.covariance(th)
    p.potrf()
    Lg = p.get_matrix()
    A = orc.covariance_matrix(0, (x, y), (x, y), th, dts=dts, lower=False)
    Lo = np.tril(orc.potrf(A, dts))
    err = np.max(np.abs(Lg - Lo))
    log("  max |L_gpu - L_oracle| = %.3e" % err)
    assert err < 1e-11
    assert np.max(np.abs(Lg @ Lg.T - A)) < 1e-12


def test_phases_must_run_in_order(ctx):
    """egs_potrf factors the matrix of the last egs_covariance call once;
    egs_solve_logdet_quad needs a factor.  Out of order, both refuse instead of
    working on stale tiles, and the in-order sequence still matches the
    one-call evaluation."""
    x, y = _rng_locs(300, 3)
    th = [1.0, 0.1, 0.5]
    p = egs.Problem(ctx, 0, x, y, dts=64)
    p.set_observations(np.linspace(-1.0, 1.0, 300))
    with pytest.raises(egs.EGSError, match="call egs_covariance first"):
        p.potrf()
    with pytest.raises(egs.EGSError, match="call egs_potrf first"):
        p.solve_logdet_quad()
    p.covariance(th)
    with pytest.raises(egs.EGSError, match="call egs_potrf first"):
        p.solve_logdet_quad()
    p.potrf()
    with pytest.raises(egs.EGSError, match="call egs_covariance first"):
        p.potrf()
    ld, q = p.solve_logdet_quad()
    _, ld1, q1 = p.loglik(th)
    assert ld == ld1 and q == q1
    p.close()


def test_not_positive_definite_reports(ctx):
    log("not-SPD error path")
    x, y = _rng_locs(100, 7)
    p = egs.Problem(ctx, 0, x, y, dts=64)
    p.set_observations(np.ones(100))
    with pytest.raises(egs.NotPositiveDefinite, match="not positive definite"):
        p.loglik([-1.0, 0.1, 0.5])
    # the problem stays usable
    llh, _, _ = p.loglik([1.0, 0.1, 0.5])
    assert np.isfinite(llh)


@pytest.mark.parametrize("mixed", [False, True])
def test_not_positive_definite_names_the_oracles_leading_minor(ctx, mixed):
    """The failing column reported by the diagonal-block kernel is the one
    dpotrf reports: sigma^2 < 0 fails at the first pivot, a bivariate cross
    correlation beta > 1 (nu1 = nu2) at the second (the first location's 2 x 2
    block has determinant 1 - beta^2 < 0)."""
    x, y = orc.generate_locations(300, seed=5)
    for kernel, th in ((0, [-1.0, 0.1, 0.5]), (1, [1.0, 1.0, 0.1, 0.5, 0.5, 1.5])):
        A = orc.covariance_matrix(kernel, (x, y), (x, y), th, dts=192, lower=False)
        with pytest.raises(RuntimeError) as info:
            orc.potrf(A, 192)
        minor = int(str(info.value).split("minor ")[1].rstrip(")"))
        log("  kernel %d: oracle leading minor %d" % (kernel, minor))
        p = egs.Problem(ctx, kernel, x, y, dts=192)
        p.set_observations(np.ones(len(x) * (2 if kernel == 1 else 1)))
        if mixed:
            p.set_precision(egs.PRECISION_MIXED, 1, 3)
        with pytest.raises(egs.NotPositiveDefinite, match=r"leading minor %d\)" % minor):
            p.loglik(th)
        p.close()


# ------------------------------------------------------------------ likelihood
@pytest.mark.parametrize("n,dts,theta", [(1600, 320, [1.0, 0.1, 0.5]), (900, 64, [1.0, 0.1, 1.0]),
                                         (1500, 960, [0.8, 0.03, 1.3]), (333, 128, [1.2, 0.2, 2.5])])
def test_loglik_matches_oracle(ctx, n, dts, theta):
    log("loglik n=%d dts=%d theta=%s" % (n, dts, theta))
    x, y = orc.generate_locations(n, seed=0)
    z = orc.generate_observations(0, (x, y), theta, dts, seed=0)
    ref = orc.loglik(0, (x, y), theta, z, dts)
    p = egs.Problem(ctx, 0, x, y, dts=dts)
    p.set_observations(z)
    llh, ld, q = p.loglik(theta)
    log("  llh gpu %.15e oracle %.15e" % (llh, ref["llh"]))
    assert abs(llh - ref["llh"]) <= 1e-10 * abs(ref["llh"])
    assert abs(ld - ref["logdet"]) <= 1e-10 * max(1.0, abs(ref["logdet"]))
    assert abs(q - ref["quad"]) <= 1e-10 * abs(ref["quad"])


def test_loglik_bivariate_matches_oracle(ctx):
    log("loglik bivariate")
    x, y = orc.generate_locations(400, seed=2)
    th = [1.0, 1.0, 0.1, 0.5, 1.0, 0.5]
    z = orc.generate_observations(1, (x, y), th, 128, seed=0)
    ref = orc.loglik(1, (x, y), th, z, 128)
    p = egs.Problem(ctx, 1, x, y, dts=128)
    p.set_observations(z)
    llh, _, _ = p.loglik(th)
    assert abs(llh - ref["llh"]) <= 1e-10 * abs(ref["llh"])


def sphere_locations(n, seed):
    """Uniform points on the sphere as (latitude, longitude) in degrees."""
    rng = np.random.default_rng(seed)
    u, v = rng.random(n), rng.random(n)
    return np.degrees(np.arcsin(2.0 * u - 1.0)), 360.0 * v - 180.0


@pytest.mark.parametrize("mixed", [False, True])
def test_loglik_bivariate_great_circle_matches_oracle(ctx, mixed):
    """The bivariate parsimonious Matérn with haversine distances (km, range
    637.1 km = 0.1 earth radii), the workload of the 25k-location case at a
    size the oracle factors in seconds; the mixed plan within 1e-6.  The
    oracle's likelihood builds Sigma with metric 0 as the reference's MLE does,
    so the expected value comes from the oracle's great-circle covariance and
    LAPACK (Cholesky, log-determinant, solve)."""
    log("loglik bivariate great-circle mixed=%s" % mixed)
    lat, lon = sphere_locations(960, 4)
    th = [1.0, 1.0, 637.1, 0.5, 1.0, 0.5]
    z = orc.generate_observations(1, (lat, lon), th, 192, seed=3, metric=1)
    S = orc.covariance_matrix(1, (lat, lon), (lat, lon), th, dts=192, metric=1, lower=False)
    Lc = np.linalg.cholesky(S)
    w = np.linalg.solve(Lc, z)
    n = len(z)
    ref = -0.5 * (w @ w) - np.sum(np.log(np.diag(Lc))) - 0.5 * n * np.log(2.0 * np.pi)
    p = egs.Problem(ctx, 1, lat, lon, metric=1, dts=192)
    p.set_observations(z)
    if mixed:
        p.set_precision(egs.PRECISION_MIXED, 1, 1)
    llh, _, _ = p.loglik(th)
    tol = 1e-6 if mixed else 1e-10
    assert abs(llh - ref) <= tol * abs(ref), (llh, ref)


def test_loglik_fixture_first_bobyqa_points(ctx, fixtures):
    log("loglik fixture -24.026")
    f = fixtures["modeling_exact"]
    L16 = fixtures["locations16"]
    p = egs.Problem(ctx, 0, L16["x"], L16["y"], dts=f["dts"])
    p.set_observations(L16["z"])
    lb = np.array(f["lb"])
    step = 0.25 * (np.array(f["ub"]) - lb)
    vals = [p.loglik(t)[0] for t in (lb, lb + step * [1, 0, 0], lb + step * [0, 1, 0])]
    assert abs(max(vals) - f["loglik"]) < f["tol"]


# ------------------------------------------------------------------ synthetic data
@pytest.mark.parametrize("key,kernel", [("z_uni_n9", 0), ("z_uni_n16", 0), ("z_biv_n16", 1)])
def test_generate_observations_fixture(ctx, fixtures, key, kernel):
    log("generate observations %s" % key)
    f = fixtures[key]
    x, y = egs.generate_locations(f["n"], seed=f["seed"])
    p = egs.Problem(ctx, kernel, x, y, dts=f["dts"])
    z = p.generate_observations(f["theta"], seed=f["seed"])
    assert np.max(np.abs(z[:len(f["z"])] - f["z"])) < f["tol"]
    zo = orc.generate_observations(kernel, (x, y), f["theta"], f["dts"], seed=f["seed"])
    assert np.max(np.abs(z - zo)) < 1e-13


def test_generate_observations_large(ctx):
    log("generate observations n=3000")
    x, y = egs.generate_locations(3000, seed=4)
    th = [1.0, 0.1, 0.5]
    p = egs.Problem(ctx, 0, x, y, dts=320)
    z = p.generate_observations(th, seed=3)
    zo = orc.generate_observations(0, (x, y), th, 320, seed=3)
    assert np.max(np.abs(z - zo)) < 1e-11


# ------------------------------------------------------------------ kriging
def test_kriging_fixture(ctx, fixtures):
    log("kriging fixture")
    f = fixtures["kriging_r"]
    L16 = fixtures["locations16"]
    x, y, z = np.array(L16["x"]), np.array(L16["y"]), np.array(L16["z"])
    p = egs.Problem(ctx, 0, x, y, dts=f["dts"])
    p.set_observations(z)
    zm, _ = p.predict(f["theta"], x[:2], y[:2])
    assert np.max(np.abs(zm - z[:2])) < f["interp_tol"]
    keep = len(x) - len(f["test_x"])
    p2 = egs.Problem(ctx, 0, x[:keep], y[:keep], dts=f["dts"])
    p2.set_observations(z[:keep])
    zm, _ = p2.predict(f["theta"], f["test_x"], f["test_y"])
    assert np.max(np.abs(zm - f["expected"])) < f["tol"]


def test_mspe_fixture(ctx, fixtures):
    log("mspe fixture")
    f = fixtures["mspe"]
    L16 = fixtures["locations16"]
    z, x, y = orc.shuffle(orc.GlibcRand(1), L16["z"], L16["x"], L16["y"])
    k = f["zmiss"]
    p = egs.Problem(ctx, 0, x[k:], y[k:], dts=f["dts"])
    p.set_observations(z[k:])
    _, mspe = p.predict(f["theta"], x[:k], y[:k], z_actual=z[:k])
    assert abs(mspe - f["mspe"]) < f["rel_tol"] * f["mspe"]


def test_kriging_at_40k_interpolates_and_is_linear(ctx):
    """At the size of the reference's 40k configuration (200 x 200 grid, dts
    960), where the oracle is too slow to run: kriging at observed locations
    returns the observations (the interpolation identity), predictions scale
    with the observations, and z' Sigma^-1 z scales with their square."""
    n = 40000
    th = [1.0, 0.1, 0.5]
    x, y = egs.generate_locations(n, seed=0)
    p = egs.Problem(ctx, 0, x, y, dts=960)
    z = p.generate_observations(th, seed=0)
    log("kriging n=40000: interpolation identity")
    idx = np.arange(0, n, n // 64)[:64]
    zm, _ = p.predict(th, x[idx], y[idx])
    err = np.max(np.abs(zm - z[idx]))
    log("  max |z_pred - z| at observed points %.3e (max |z| %.3f)" % (err, np.max(np.abs(z))))
    assert err < 1e-8 * np.max(np.abs(z))
    rng = np.random.default_rng(1)
    xm, ym = rng.random(100), rng.random(100)
    z1, _ = p.predict(th, xm, ym)
    _, _, q1 = p.loglik(th)
    p.set_observations(4.0 * z)
    z4, _ = p.predict(th, xm, ym)
    _, _, q4 = p.loglik(th)
    assert np.max(np.abs(z4 - 4.0 * z1)) <= 1e-12 * np.max(np.abs(z4))
    assert abs(q4 - 16.0 * q1) <= 1e-12 * abs(q4)
    p.close()


@pytest.mark.parametrize("n_obs,dts,theta", [(2000, 320, [1.0, 0.1, 1.0]), (1100, 960, [0.9, 0.05, 0.5])])
def test_kriging_matches_oracle(ctx, n_obs, dts, theta):
    log("kriging n_obs=%d" % n_obs)
    x, y = orc.generate_locations(n_obs, seed=9)
    z = orc.generate_observations(0, (x, y), theta, dts, seed=1)
    rng = np.random.default_rng(0)
    xm, ym = rng.random(50), rng.random(50)
    p = egs.Problem(ctx, 0, x, y, dts=dts)
    p.set_observations(z)
    zm, _ = p.predict(theta, xm, ym)
    zo, _ = orc.krige(0, (x, y), (xm, ym), theta, z, dts)
    assert np.max(np.abs(zm - zo)) < 1e-8


@pytest.mark.parametrize("mixed", [False, True])
def test_bivariate_kriging_matches_oracle(ctx, mixed):
    """Bivariate kriging with the reference's three errors: predictions within
    1e-8 of the oracle, [mspe, mspe1, mspe2] within 1e-8 relative (FP64 plan;
    the mixed plan's refined solve within 1e-7 on the predictions)."""
    log("bivariate kriging mixed=%s" % mixed)
    theta = [1.0, 1.0, 0.1, 0.5, 1.0, 0.5]
    x, y = orc.generate_locations(700, seed=21)
    z = orc.generate_observations(1, (x, y), theta, 192, seed=2)
    rng = np.random.default_rng(5)
    xm, ym = rng.random(40), rng.random(40)
    za = rng.standard_normal(80)
    p = egs.Problem(ctx, 1, x, y, dts=192)
    p.set_observations(z)
    if mixed:
        p.set_precision(egs.PRECISION_MIXED, 1, 1)
    zm, m3 = p.predict_mspe(theta, xm, ym, z_actual=za)
    zo, mo = orc.krige(1, (x, y), (xm, ym), theta, z, 192, z_actual=za)
    mo3 = orc.mspe(1, za, zo, 192)
    assert abs(mo3[0] - mo) <= 1e-14 * mo
    tol = 1e-7 if mixed else 1e-8
    assert np.max(np.abs(zm - zo)) < tol
    assert np.allclose(m3, mo3, rtol=tol, atol=0)
    # the pooled error through egs_predict is the first of the three
    _, mspe = p.predict(theta, xm, ym, z_actual=za)
    assert mspe == m3[0]
    p.close()


def test_univariate_predict_mspe_leaves_per_variable_errors_zero(ctx):
    theta = [1.0, 0.1, 1.0]
    x, y = orc.generate_locations(400, seed=4)
    z = orc.generate_observations(0, (x, y), theta, 128, seed=3)
    rng = np.random.default_rng(8)
    xm, ym, za = rng.random(30), rng.random(30), rng.standard_normal(30)
    p = egs.Problem(ctx, 0, x, y, dts=128)
    p.set_observations(z)
    zm, m3 = p.predict_mspe(theta, xm, ym, z_actual=za)
    zo, mo = orc.krige(0, (x, y), (xm, ym), theta, z, 128, z_actual=za)
    assert np.max(np.abs(zm - zo)) < 1e-8
    assert abs(m3[0] - mo) <= 1e-8 * mo and m3[1] == 0.0 and m3[2] == 0.0
    _, m_none = p.predict_mspe(theta, xm, ym)
    assert list(m_none) == [-1.0, 0.0, 0.0]
    p.close()


@pytest.mark.parametrize("kernel,theta", [(0, [1.0, 0.1, 0.5]), (1, [1.0, 1.0, 0.1, 0.5, 1.0, 0.5])])
def test_covariance_tile_interior_offsets_match_oracle(ctx, kernel, theta):
    """A tile away from the origin of both location sets (the plugin stages only
    the locations it touches): equal to the oracle's tile at 1e-13 relative,
    including ragged tiles.  Bivariate tiles have even row and column counts,
    as every tile of the reference's even dense tile size does (for an odd row
    count its loop writes row i+1 of the last pair into the next column,
    BivariateMaternParsimonious.cpp:80-95)."""
    x, y = orc.generate_locations(900, seed=6)
    xs, ys = x[::-1].copy(), y[::-1].copy()
    ragged = ((18, 34, 1000, 12), (6, 2, 0, 0)) if kernel else ((17, 33, 500, 12), (5, 3, 0, 0))
    for rows, cols, r0, c0 in ((96, 64, 320, 640),) + ragged:
        got = ctx.covariance_tile(kernel, rows, cols, r0, c0, (x, y), (xs, ys), theta)
        ref = orc.covariance_tile(kernel, rows, cols, r0, c0, (x, y), (xs, ys), theta)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
