"""Pins the CPU oracle (oracle/egs_oracle.c) before it is trusted as the checker.

Each oracle function is checked against the reference's own fixtures
(tests/golden/reference_fixtures.json, data copied from the reference's tests)
and against what the mathematics fixes: libc's rand(), 30-digit mpmath Bessel
values, closed forms, numpy/scipy LAPACK.
"""
import ctypes
import ctypes.util
import math
import os

import numpy as np
import pytest

from oracle import orc


def _libc():
    return ctypes.CDLL(ctypes.util.find_library("c") or "libc.so.6")


# ---------------------------------------------------------------- random numbers
@pytest.mark.parametrize("seed", [0, 1, 7, 42, 123456789])
def test_glibc_rand_matches_libc(seed):
    libc = _libc()
    libc.srand(ctypes.c_uint(seed))
    expect = [libc.rand() for _ in range(2000)]
    r = orc.GlibcRand(seed)
    assert [r.rand() for _ in range(2000)] == expect


def test_dlaruv_is_one_lcg_stream():
    a = 33952834046453
    u, seed = orc.dlaruv([0, 0, 0, 1], 3)
    m = 1 << 48
    s = 1
    for k in range(3):
        s = (s * a) % m
        assert u[k] == s / m
    # two calls of 64 continue the same stream as one call of 128
    u128, s128 = orc.dlaruv([3, 5, 7, 9], 128)
    u1, s1 = orc.dlaruv([3, 5, 7, 9], 64)
    u2, s2 = orc.dlaruv(s1, 64)
    assert np.array_equal(u128, np.concatenate([u1, u2])) and s2 == s128


def test_dlarnv_first_value_matches_fixture(fixtures):
    # z[0] = L00 * e0 and L00 = sqrt(sigma^2) = 1, so e0 is the fixture's z[0]
    e, _ = orc.dlarnv_normal([0, 0, 0, 1], 16)
    assert abs(e[0] - fixtures["locations16"]["z"][0]) < 1e-15


def test_dlarnv_matches_mkl_when_present():
    path = "/opt/conda/lib/libmkl_rt.so.1"
    if not os.path.exists(path):
        pytest.skip("MKL not present")
    mkl = ctypes.CDLL(path)
    for n in (1, 63, 64, 65, 1000):
        iseed = (ctypes.c_int * 4)(5, 17, 99, 1)
        x = np.empty(n)
        mkl.dlarnv_(ctypes.byref(ctypes.c_int(3)), iseed, ctypes.byref(ctypes.c_int(n)),
                    x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        y, s = orc.dlarnv_normal([5, 17, 99, 1], n)
        assert np.max(np.abs(x - y)) < 1e-14
        assert list(iseed) == s


# ---------------------------------------------------------------- locations
def test_morton_constants(fixtures):
    m = fixtures["morton"]
    assert orc.spread_bits(0x7FFF) == int(m["spread_7fff"], 16)
    assert orc.reverse_spread_bits(int(m["spread_7fff"], 16)) == 0x7FFF
    s = orc.spread_bits(0x7FFF)
    assert (s << 2) == int(m["z_shift2_7fff"], 16)
    assert (s << 1) == int(m["y_shift1_7fff"], 16)
    assert (s << 2) + (s << 1) + s == int(m["xyz_7fff"], 16)
    assert (s << 1) + s == int(m["yz_7fff_z0"], 16)
    x, y, z = m["roundtrip"]
    v = (orc.spread_bits(z) << 2) + (orc.spread_bits(y) << 1) + orc.spread_bits(x)
    assert (orc.reverse_spread_bits(v), orc.reverse_spread_bits(v >> 1), orc.reverse_spread_bits(v >> 2)) == (x, y, z)
    for b in range(16):
        assert orc.spread_bits(1 << b) == 1 << (4 * b)


def test_locations_n9_fixture(fixtures):
    f = fixtures["locations9"]
    x, y = orc.generate_locations(9, seed=0)
    assert np.max(np.abs(x - f["x"])) < f["tol"]
    assert np.max(np.abs(y - f["y"])) < f["tol"]
    # re-generation from the same stream gives new points, re-seeding the same ones
    rng = orc.GlibcRand(0)
    orc.generate_locations(9, rng=rng)
    x2, _ = orc.generate_locations(9, rng=rng)
    assert np.all(np.abs(x2 - f["x"]) > 1e-6) or not np.allclose(x2, f["x"])


def test_locations_n16_exact(fixtures):
    f = fixtures["locations16"]
    x, y = orc.generate_locations(16, seed=0)
    assert np.array_equal(x, np.array(f["x"]))
    assert np.array_equal(y, np.array(f["y"]))


def test_locations_snap_to_16bit_grid():
    x, y = orc.generate_locations(400, seed=3)
    assert np.array_equal(np.round(x * 65535) / 65535, x)
    assert np.all((x > 0) & (x < 1) & (y > 0) & (y < 1))


# ---------------------------------------------------------------- Bessel K_nu
@pytest.mark.parametrize("nu", [0.01, 0.3, 0.5, 0.75, 1.0, 1.234, 1.5, 2.0, 2.5, 3.7, 5.0])
def test_bessel_knu_vs_mpmath(nu):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    for x in (1e-9, 1e-4, 0.05, 0.5, 1.0, 1.999, 2.0, 2.001, 3.0, 7.5, 20.0, 80.0, 300.0, 700.0):
        ref = float(mpmath.besselk(nu, x))
        v = orc.bessel_knu(nu, x)
        assert abs(v - ref) <= 1e-14 * abs(ref), (nu, x, v, ref)


def test_bessel_half_integer_closed_forms():
    for x in np.geomspace(1e-6, 50, 60):
        k05 = math.sqrt(math.pi / (2 * x)) * math.exp(-x)
        assert abs(orc.bessel_knu(0.5, x) - k05) <= 2e-15 * k05
        k15 = k05 * (1 + 1 / x)
        assert abs(orc.bessel_knu(1.5, x) - k15) <= 2e-15 * k15
        k25 = k05 * (1 + 3 / x + 3 / x ** 2)
        assert abs(orc.bessel_knu(2.5, x) - k25) <= 3e-15 * k25


def test_bessel_vs_scipy():
    sp = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(5)
    nu = rng.uniform(0.05, 4.5, 400)
    x = np.exp(rng.uniform(np.log(1e-3), np.log(60), 400))
    ours = np.array([orc.bessel_knu(a, b) for a, b in zip(nu, x)])
    assert np.max(np.abs(ours - sp.kv(nu, x)) / sp.kv(nu, x)) < 1e-13


# ---------------------------------------------------------------- distances
def test_distances(fixtures):
    d = fixtures["distance"]
    loc = (np.array([10.0, 4.0]), np.array([12.0, 4.0]))
    A = orc.covariance_tile(0, 1, 1, 0, 1, loc, loc, [1.0, 1.0, 0.5])
    assert abs(A[0, 0] - math.exp(-d["euclid"])) < 1e-15
    assert orc.distance_earth(0.1, 0.2, 0.1, 0.2) == 0.0
    assert abs(orc.distance_earth(90, 0, -90, 0) - math.pi * 6371.0) < 1e-9
    assert abs(orc.distance_earth(0, 0, 0, 90) - math.pi * 6371.0 / 2) < 1e-9
    ny = orc.distance_earth(*d["ny_la"])
    assert abs(ny - d["ny_la_km"]) <= d["ny_la_rel"] * d["ny_la_km"]
    # metric 1 in a tile uses the haversine with x = lat, y = lon
    A1 = orc.covariance_tile(0, 1, 1, 0, 1, (np.array([40.7128, 34.0522]), np.array([-74.0060, -118.2437])),
                             (np.array([40.7128, 34.0522]), np.array([-74.0060, -118.2437])), [1.0, 1000.0, 0.5],
                             metric=1)
    assert abs(A1[0, 0] - math.exp(-ny / 1000.0)) < 1e-14
    # great-circle distance with a third coordinate is rejected
    loc3 = (np.array([10.0, 4.0]), np.array([12.0, 4.0]), np.array([1.0, 2.0]))
    with pytest.raises(RuntimeError):
        orc.covariance_tile(0, 1, 1, 0, 1, loc3, loc3, [1.0, 1.0, 0.5], metric=1)
    A3 = orc.covariance_tile(0, 1, 1, 0, 1, loc3, loc3, [1.0, 1.0, 0.5])
    d3 = -math.log(A3[0, 0])
    assert abs(d3 - d["euclid3d"]) <= d["euclid3d_rel"] * d["euclid3d"]


# ---------------------------------------------------------------- covariance
def test_matern_closed_forms_and_offsets():
    rng = np.random.default_rng(1)
    x, y = rng.random(50), rng.random(50)
    D = np.hypot(x[:, None] - x[None, :], y[:, None] - y[None, :])
    th = [1.7, 0.13, 0.5]
    C = orc.covariance_matrix(0, (x, y), (x, y), th, dts=16, lower=False)
    assert np.allclose(C, 1.7 * np.exp(-D / 0.13), rtol=1e-14, atol=0)
    th = [0.6, 0.2, 1.5]
    C = orc.covariance_matrix(0, (x, y), (x, y), th, dts=7, lower=False)
    Dn = D / 0.2
    assert np.allclose(C, 0.6 * (1 + Dn) * np.exp(-Dn), rtol=1e-14, atol=0)
    # lower-only build leaves strictly-upper tiles untouched (zero) and matches the full build below
    L = orc.covariance_matrix(0, (x, y), (x, y), th, dts=7, lower=True)
    assert np.array_equal(np.tril(L), np.tril(C))
    assert np.all(L[:7, 7:] == 0)
    # a tile at an offset equals the same block of the full matrix
    T = orc.covariance_tile(0, 5, 9, 11, 3, (x, y), (x, y), th)
    assert np.array_equal(T, C[11:16, 3:12])


def test_matern_general_nu_vs_scipy():
    sp = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(2)
    x, y = rng.random(40), rng.random(40)
    th = [1.3, 0.07, 1.234]
    C = orc.covariance_matrix(0, (x, y), (x, y), th, dts=13, lower=False)
    D = np.hypot(x[:, None] - x[None, :], y[:, None] - y[None, :]) / th[1]
    with np.errstate(invalid="ignore", divide="ignore"):
        R = th[0] / (2 ** (th[2] - 1) * sp.gamma(th[2])) * D ** th[2] * sp.kv(th[2], D)
    R[D == 0] = th[0]
    assert np.allclose(C, R, rtol=1e-13, atol=0)


def test_bivariate_structure():
    rng = np.random.default_rng(3)
    x, y = rng.random(12), rng.random(12)
    th = [1.0, 2.0, 0.1, 0.5, 1.5, 0.4]
    C = orc.covariance_matrix(1, (x, y), (x, y), th, dts=6, lower=False)
    assert C.shape == (24, 24)
    assert np.allclose(C, C.T, rtol=0, atol=1e-15)
    # var1 block is Matern(nu1), var2 block is Matern(nu2) with their sigma^2
    D = np.hypot(x[:, None] - x[None, :], y[:, None] - y[None, :]) / 0.1
    assert np.allclose(C[0::2, 0::2], 1.0 * np.exp(-D), rtol=1e-14)
    assert np.allclose(C[1::2, 1::2], 2.0 * (1 + D) * np.exp(-D), rtol=1e-14)
    nu12 = 1.0
    rho = 0.4 * math.sqrt(math.gamma(1.5) * math.gamma(2.5) / (math.gamma(0.5) * math.gamma(1.5))) \
        * math.gamma(nu12) / math.gamma(nu12 + 1)
    assert abs(C[0, 1] - rho * math.sqrt(2.0)) < 1e-15
    np.linalg.cholesky(C)


# ---------------------------------------------------------------- dense linear algebra
def _spd(n, seed):
    rng = np.random.default_rng(seed)
    x, y = rng.random(n), rng.random(n)
    return orc.covariance_matrix(0, (x, y), (x, y), [1.0, 0.1, 0.5], dts=n, lower=False)


@pytest.mark.parametrize("n,nb", [(1, 4), (37, 8), (100, 32), (257, 64), (300, 300)])
def test_potrf_vs_numpy(n, nb):
    A = _spd(n, n)
    L = np.tril(orc.potrf(A, nb))
    R = np.linalg.cholesky(A)
    assert np.allclose(L, R, rtol=0, atol=1e-12)
    assert np.allclose(L @ L.T, A, rtol=0, atol=1e-13)


def test_potrf_not_spd_reports():
    A = _spd(20, 4)
    A[5, 5] = -1.0
    with pytest.raises(RuntimeError, match="not positive definite"):
        orc.potrf(A, 8)


def test_logdet_trsv_trmv():
    lib = orc.lib()
    A = _spd(90, 9)
    L = np.asfortranarray(np.tril(orc.potrf(A, 16)))
    dp = ctypes.POINTER(ctypes.c_double)
    half = lib.orc_logdet_half(L.ctypes.data_as(dp), 90)
    assert abs(2 * half - np.linalg.slogdet(A)[1]) < 1e-11
    rng = np.random.default_rng(0)
    b = rng.standard_normal(90)
    z = b.copy()
    lib.orc_trsv_lower(L.ctypes.data_as(dp), 90, z.ctypes.data_as(dp))
    assert np.allclose(L @ z, b, atol=1e-12)
    w = z.copy()
    lib.orc_trsv_lower_trans(L.ctypes.data_as(dp), 90, w.ctypes.data_as(dp))
    assert np.allclose(A @ w, b, atol=1e-10)
    v = b.copy()
    lib.orc_trmv_lower(L.ctypes.data_as(dp), 90, v.ctypes.data_as(dp))
    assert np.allclose(v, L @ b, atol=1e-13)


# ---------------------------------------------------------------- the hot path
@pytest.mark.parametrize("key", ["z_uni_n9", "z_uni_n16"])
def test_observations_fixture_univariate(fixtures, key):
    f = fixtures[key]
    x, y = orc.generate_locations(f["n"], seed=f["seed"])
    z = orc.generate_observations(0, (x, y), f["theta"], f["dts"], seed=f["seed"])
    assert np.max(np.abs(z - f["z"])) < f["tol"]


def test_observations_fixture_full_precision(fixtures):
    f = fixtures["locations16"]
    x, y = orc.generate_locations(16, seed=0)
    z = orc.generate_observations(0, (x, y), [1, 0.1, 0.5], 9, seed=0)
    assert np.max(np.abs(z - f["z"])) < 4e-15


def test_observations_fixture_bivariate(fixtures):
    f = fixtures["z_biv_n16"]
    x, y = orc.generate_locations(f["n"], seed=f["seed"])
    z = orc.generate_observations(1, (x, y), f["theta"], f["dts"], seed=f["seed"])
    assert len(z) == 32
    assert np.max(np.abs(z[:16] - f["z"])) < f["tol"]


def test_loglik_fixture_bobyqa_first_points(fixtures):
    """The EXACT_DENSE fixture is the best of BOBYQA's first three evaluations from lb with NLopt's default
    initial step 0.25 (ub - lb) = 1.225: lb, lb + 1.225 e1, lb + 1.225 e2."""
    f = fixtures["modeling_exact"]
    loc = (np.array(fixtures["locations16"]["x"]), np.array(fixtures["locations16"]["y"]))
    z = np.array(fixtures["locations16"]["z"])
    lb = np.array(f["lb"])
    step = 0.25 * (np.array(f["ub"]) - lb)
    pts = [lb, lb + step * [1, 0, 0], lb + step * [0, 1, 0]]
    vals = [orc.loglik(0, loc, p, z, f["dts"])["llh"] for p in pts]
    assert abs(max(vals) - f["loglik"]) < f["tol"]


def test_loglik_matches_gaussian_density():
    sp = pytest.importorskip("scipy.stats")
    x, y = orc.generate_locations(200, seed=11)
    th = [1.3, 0.08, 0.9]
    rng = np.random.default_rng(0)
    z = rng.standard_normal(200)
    r = orc.loglik(0, (x, y), th, z, 37)
    C = orc.covariance_matrix(0, (x, y), (x, y), th, dts=37, lower=False)
    ref = sp.multivariate_normal(mean=np.zeros(200), cov=C).logpdf(z)
    assert abs(r["llh"] - ref) < 1e-9 * abs(ref)
    assert abs(r["quad"] - z @ np.linalg.solve(C, z)) < 1e-8 * r["quad"]
    assert abs(r["logdet"] - np.linalg.slogdet(C)[1]) < 1e-9


def test_kriging_fixture_and_interpolation(fixtures):
    f = fixtures["kriging_r"]
    L16 = fixtures["locations16"]
    x, y, z = np.array(L16["x"]), np.array(L16["y"]), np.array(L16["z"])
    # predicting at training points returns the training values
    zm, _ = orc.krige(0, (x, y), (x[:2], y[:2]), f["theta"], z, f["dts"])
    assert np.max(np.abs(zm - z[:2])) < f["interp_tol"]
    # the reference's explicit train/test path keeps n_obs = n_train - n_test observations
    n_keep = len(x) - len(f["test_x"])
    zm, _ = orc.krige(0, (x[:n_keep], y[:n_keep]), (np.array(f["test_x"]), np.array(f["test_y"])), f["theta"],
                      z[:n_keep], f["dts"])
    assert np.max(np.abs(zm - f["expected"])) < f["tol"]


def test_mspe_fixture(fixtures):
    f = fixtures["mspe"]
    L16 = fixtures["locations16"]
    rng = orc.GlibcRand(1)  # an unseeded glibc rand() starts as srand(1)
    z, x, y = orc.shuffle(rng, L16["z"], L16["x"], L16["y"])
    k = f["zmiss"]
    _, mspe = orc.krige(0, (x[k:], y[k:]), (x[:k], y[:k]), f["theta"], z[k:], f["dts"], z_actual=z[:k])
    assert abs(mspe - f["mspe"]) < f["rel_tol"] * f["mspe"]


def test_kriging_vs_numpy():
    rng = np.random.default_rng(4)
    xo, yo = rng.random(120), rng.random(120)
    xm, ym = rng.random(7), rng.random(7)
    th = [0.8, 0.15, 1.0]
    zo = rng.standard_normal(120)
    zm, _ = orc.krige(0, (xo, yo), (xm, ym), th, zo, 25)
    C22 = orc.covariance_matrix(0, (xo, yo), (xo, yo), th, dts=120, lower=False)
    C12 = orc.covariance_matrix(0, (xm, ym), (xo, yo), th, dts=120, lower=False)
    assert np.allclose(zm, C12 @ np.linalg.solve(C22, zo), atol=1e-10)


def test_mspe_per_variable_hand_computed():
    """[mspe, mspe1, mspe2] of LinearAlgebraMethods.cpp:575-600: errors
    1..6 in tiles of 4 give squared sums 91 (all), 1+9+25 (even entries, the
    first variable) and 4+16+36 (odd entries), over 6, 3 and 3 values."""
    out = orc.mspe(1, [1, 2, 3, 4, 5, 6], np.zeros(6), 4)
    assert np.allclose(out, [91 / 6, 35 / 3, 56 / 3], rtol=0, atol=1e-15)
    # univariate: the per-variable errors stay 0 as the reference leaves them
    assert np.allclose(orc.mspe(0, [1, 2, 3], np.zeros(3), 2), [14 / 3, 0, 0], rtol=0, atol=1e-15)


def test_mspe_matches_krige_and_splits_evenly():
    """The pooled error equals orc_krige's and the mean of the two per-variable
    errors (an equal number of values of each variable); the tile size does not
    change the result beyond summation order."""
    rng = np.random.default_rng(11)
    xo, yo = rng.random(60), rng.random(60)
    xm, ym = rng.random(9), rng.random(9)
    th = [1.0, 1.0, 0.1, 0.5, 1.0, 0.5]
    zo = rng.standard_normal(120)
    za = rng.standard_normal(18)
    zm, mspe = orc.krige(1, (xo, yo), (xm, ym), th, zo, 20, z_actual=za)
    out = orc.mspe(1, za, zm, 20)
    assert abs(out[0] - mspe) <= 1e-14 * mspe
    assert abs(out[0] - 0.5 * (out[1] + out[2])) <= 1e-14 * out[0]
    assert np.allclose(out, orc.mspe(1, za, zm, 6), rtol=1e-14, atol=0)
    d = (za - zm) ** 2
    assert np.allclose(out, [d.mean(), d[0::2].mean(), d[1::2].mean()], rtol=1e-14, atol=0)
