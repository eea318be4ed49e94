"""ctypes wrapper of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Imported only by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Every
function forwards to the C restatement in egs_oracle.c, which cites the
reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_bessel_knu.restype = ctypes.c_double
        L.orc_bessel_knu.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_distance_earth.restype = ctypes.c_double
        L.orc_distance_earth.argtypes = [ctypes.c_double] * 4
        L.orc_spread_bits.restype = ctypes.c_uint64
        L.orc_spread_bits.argtypes = [ctypes.c_uint64]
        L.orc_reverse_spread_bits.restype = ctypes.c_uint64
        L.orc_reverse_spread_bits.argtypes = [ctypes.c_uint64]
        L.orc_rand.restype = ctypes.c_int
        L.orc_potrf.restype = _i64
        L.orc_potrf.argtypes = [_dp, _i64, ctypes.c_int]
        L.orc_logdet_half.restype = ctypes.c_double
        L.orc_logdet_half.argtypes = [_dp, _i64]
        L.orc_dot.restype = ctypes.c_double
        for name in ("orc_trsv_lower", "orc_trsv_lower_trans", "orc_trmv_lower"):
            getattr(L, name).argtypes = [_dp, _i64, _dp]
        L.orc_covariance_matrix.argtypes = [ctypes.c_int, _dp, _i64, _i64, ctypes.c_int, ctypes.c_int,
                                            _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_covariance_tile.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_loglik.argtypes = [ctypes.c_int, _i64, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp,
                                 _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_generate_observations.argtypes = [ctypes.c_int, _i64, ctypes.c_int, _dp, _dp, _dp, _dp,
                                                ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_mspe.argtypes = [ctypes.c_int, _i64, ctypes.c_int, _dp, _dp, _dp]
        L.orc_krige.argtypes = [ctypes.c_int, _i64, _i64, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


class RandState(ctypes.Structure):
    _fields_ = [("r", ctypes.c_int32 * 34), ("f", ctypes.c_int), ("b", ctypes.c_int)]


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class GlibcRand:
    """glibc rand() restated (ExaGeoStat.cpp:76 srand, LocationGenerator.cpp:84-89)."""

    def __init__(self, seed):
        self.state = RandState()
        lib().orc_srand(ctypes.byref(self.state), ctypes.c_uint(seed))

    def rand(self):
        return lib().orc_rand(ctypes.byref(self.state))


def dlarnv_normal(seed4, n):
    iseed = (ctypes.c_int * 4)(*seed4)
    out = np.empty(n, dtype=np.float64)
    lib().orc_dlarnv_normal(iseed, _i64(n), _p(out))
    return out, list(iseed)


def dlaruv(seed4, n):
    iseed = (ctypes.c_int * 4)(*seed4)
    out = np.empty(n, dtype=np.float64)
    lib().orc_dlaruv(iseed, ctypes.c_int(n), _p(out))
    return out, list(iseed)


def spread_bits(v):
    return lib().orc_spread_bits(v)


def reverse_spread_bits(v):
    return lib().orc_reverse_spread_bits(v)


def generate_locations(n, seed=0, dim=2, rng=None):
    """LocationGenerator::GenerateLocations after srand(seed) (LocationGenerator.cpp:25-123)."""
    rng = rng if rng is not None else GlibcRand(seed)
    x = np.empty(n)
    y = np.empty(n)
    z = np.empty(n) if dim == 3 else None
    lib().orc_generate_locations(ctypes.c_int(n), ctypes.c_int(1 if dim == 3 else 0), ctypes.byref(rng.state),
                                 _p(x), _p(y), _p(z))
    return (x, y) if dim != 3 else (x, y, z)


def bessel_knu(nu, x):
    return lib().orc_bessel_knu(float(nu), float(x))


def distance_earth(lat1, lon1, lat2, lon2):
    return lib().orc_distance_earth(lat1, lon1, lat2, lon2)


KERNELS = {"UnivariateMaternStationary": 0, "univariate_matern_stationary": 0,
           "BivariateMaternParsimonious": 1, "bivariate_matern_parsimonious": 1}


def _kid(kernel):
    return KERNELS[kernel] if isinstance(kernel, str) else int(kernel)


def covariance_tile(kernel, rows, cols, row_off, col_off, loc1, loc2, theta, metric=0):
    A = np.zeros((cols, rows))  # column-major rows x cols
    x1, y1 = _f64(loc1[0]), _f64(loc1[1])
    z1 = _f64(loc1[2]) if len(loc1) > 2 else None
    x2, y2 = _f64(loc2[0]), _f64(loc2[1])
    z2 = _f64(loc2[2]) if len(loc2) > 2 else None
    th = _f64(theta)
    st = lib().orc_covariance_tile(_kid(kernel), _p(A), rows, cols, row_off, col_off, _p(x1), _p(y1), _p(z1),
                                   _p(x2), _p(y2), _p(z2), _p(th), metric)
    if st:
        raise RuntimeError("oracle covariance tile failed (%d)" % st)
    return A.T.copy()


def covariance_matrix(kernel, loc1, loc2, theta, dts, metric=0, lower=True, order="C"):
    """Sigma built tile-by-tile as DCMGCodelet::InsertTask (dcmg-codelet.cpp:47-74); returns a
    numpy (m, n) array (lower tiles only when lower and square), C-contiguous, or with
    order="F" the column-major result itself (no copy)."""
    p = 2 if _kid(kernel) == 1 else 1
    m = len(loc1[0]) * p
    n = len(loc2[0]) * p
    C = np.zeros((n, m))  # column-major m x n
    x1, y1 = _f64(loc1[0]), _f64(loc1[1])
    z1 = _f64(loc1[2]) if len(loc1) > 2 else None
    x2, y2 = _f64(loc2[0]), _f64(loc2[1])
    z2 = _f64(loc2[2]) if len(loc2) > 2 else None
    th = _f64(theta)
    st = lib().orc_covariance_matrix(_kid(kernel), _p(C), m, n, dts, 1 if lower else 0, _p(x1), _p(y1), _p(z1),
                                     _p(x2), _p(y2), _p(z2), _p(th), metric)
    if st:
        raise RuntimeError("oracle covariance failed (%d)" % st)
    return C.T if order == "F" else C.T.copy()


def potrf(A, nb=256):
    """Lower tiled Cholesky of a numpy (n, n) array; returns L (lower, upper untouched)."""
    n = A.shape[0]
    F = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    info = lib().orc_potrf(F.ctypes.data_as(_dp), n, nb)
    if info:
        raise RuntimeError("Matrix is not positive definite (minor %d)" % info)
    return F


def loglik(kernel, loc, theta, z, dts, metric=0):
    """One ExaGeoStatMLETile evaluation. Returns dict(llh, logdet, quad, z_solved, t_*)."""
    kid = _kid(kernel)
    p = 2 if kid == 1 else 1
    n_loc = len(loc[0])
    n = n_loc * p
    C = np.empty(n * n)
    zz = _f64(z).copy()
    x, y = _f64(loc[0]), _f64(loc[1])
    zc = _f64(loc[2]) if len(loc) > 2 else None
    out = [ctypes.c_double() for _ in range(6)]
    st = lib().orc_loglik(kid, n_loc, dts, _p(x), _p(y), _p(zc), _p(_f64(theta)), metric, _p(zz), _p(C),
                          *[ctypes.byref(o) for o in out])
    if st:
        raise RuntimeError("CHAMELEON_dpotrf_Tile Failed, Matrix is not positive definite")
    return {"llh": out[0].value, "logdet": out[1].value, "quad": out[2].value, "z_solved": zz,
            "t_gen": out[3].value, "t_facto": out[4].value, "t_solve": out[5].value}


def generate_observations(kernel, loc, theta, dts, seed=0, metric=0):
    kid = _kid(kernel)
    p = 2 if kid == 1 else 1
    n_loc = len(loc[0])
    n = n_loc * p
    C = np.empty(n * n)
    z = np.empty(n)
    x, y = _f64(loc[0]), _f64(loc[1])
    zc = _f64(loc[2]) if len(loc) > 2 else None
    st = lib().orc_generate_observations(kid, n_loc, dts, _p(x), _p(y), _p(zc), _p(_f64(theta)), metric, seed,
                                         _p(z), _p(C))
    if st:
        raise RuntimeError("CHAMELEON_dpotrf_Tile Failed, Matrix is not positive definite")
    return z


def krige(kernel, obs_loc, miss_loc, theta, zobs, dts, z_actual=None, metric=0):
    kid = _kid(kernel)
    p = 2 if kid == 1 else 1
    n_obs = len(obs_loc[0])
    n_miss = len(miss_loc[0])
    zmiss = np.empty(n_miss * p)
    mspe = ctypes.c_double()
    za = _f64(z_actual)
    st = lib().orc_krige(kid, n_obs, n_miss, dts, _p(_f64(obs_loc[0])), _p(_f64(obs_loc[1])),
                         _p(_f64(miss_loc[0])), _p(_f64(miss_loc[1])), _p(_f64(theta)), metric,
                         _p(_f64(zobs)), _p(za), _p(zmiss), ctypes.byref(mspe))
    if st:
        raise RuntimeError("CHAMELEON_dposv_Tile Failed!")
    return zmiss, mspe.value


def mspe(kernel, z_actual, zmiss, dts):
    """[mspe, mspe1, mspe2] as ExaGeoStatMLEPredictTile returns them
    (LinearAlgebraMethods.cpp:575-600; dmse / dmse-bivariate codelets)."""
    za, zm = _f64(z_actual), _f64(zmiss)
    out = np.zeros(3)
    lib().orc_mspe(_kid(kernel), len(za), int(dts), _p(za), _p(zm), _p(out))
    return out


def shuffle(rng, z, x, y):
    z = _f64(z).copy()
    x = _f64(x).copy()
    y = _f64(y).copy()
    lib().orc_shuffle(ctypes.byref(rng.state), _p(z), _p(x), _p(y), ctypes.c_int(len(z)))
    return z, x, y
