"""ctypes binding of the C ABI in include/exageostat_amd.h.

The library must be built (``__graft_entry__.build()`` or
``python exageostatcpp_amd/build.py``); loading fails loudly otherwise.  There
is no CPU fallback: every numerical result comes from the gfx950 kernels.
"""
import atexit
import ctypes
import ctypes.util
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexageostat_amd.so")

SUCCESS = 0
ERR_INVALID_ARG = 1
ERR_HIP = 2
ERR_NOT_SPD = 3
ERR_OUT_OF_MEMORY = 4
ERR_UNSUPPORTED = 5
ERR_NO_DEVICE = 6
ERR_STATE = 7

KERNEL_UNIVARIATE = 0
KERNEL_BIVARIATE = 1
KERNELS = {"UnivariateMaternStationary": 0, "univariate_matern_stationary": 0,
           "BivariateMaternParsimonious": 1, "bivariate_matern_parsimonious": 1}
PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_MIXED_BF16 = 2
PROFILE_UPDATE = 0
PROFILE_COVARIANCE = 1
PROFILE_PANEL_TRSM = 2
PROFILE_DIAGONAL = 3
PROFILE_SOLVE = 4
PROFILE_UPDATE_F32 = 5
PROFILE_UPDATE_16 = 6
PROFILE_REFINE = 7
PROFILE_UPDATE_16_SELECT = 8

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p
_lib = None


class EGSError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class NotPositiveDefinite(EGSError):
    pass


EXPORTS = {
    "egs_last_error": ([], ctypes.c_char_p),
    "egs_version": ([], _int),
    "egs_init": ([_int, ctypes.POINTER(_vp)], _int),
    "egs_finalize": ([_vp], _int),
    "egs_synchronize": ([_vp], _int),
    "egs_covariance_tile": ([_vp, _int, _dp, _int, _int, _int, _int, _dp, _dp, _dp, _i64, _dp, _dp, _dp, _i64, _dp,
                             _int], _int),
    "egs_covariance_tile_noalloc": ([_vp, _int, _dp, _int, _int, _int, _int, _dp, _dp, _dp, _i64, _dp, _dp, _dp, _i64,
                                     _dp, _int], _int),
    "egs_problem_create": ([_vp, _int, _i64, _dp, _dp, _dp, _int, _int, ctypes.POINTER(_vp)], _int),
    "egs_problem_destroy": ([_vp], _int),
    "egs_problem_set_context": ([_vp, _vp], _int),
    "egs_problem_device_bytes": ([_vp, ctypes.POINTER(_i64)], _int),
    "egs_problem_estimate_bytes": ([_int, _i64, _int, _int, _int, _int, _int, ctypes.POINTER(_i64)], _int),
    "egs_memory_cache": ([_i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64)], _int),
    "egs_problem_set_precision": ([_vp, _int, _int, _int], _int),
    "egs_problem_set_single_product": ([_vp, _int, ctypes.c_double], _int),
    "egs_problem_set_fp8_product": ([_vp, ctypes.c_double], _int),
    "egs_problem_set_fp64_product_band": ([_vp, _int], _int),
    "egs_problem_set_graph": ([_vp, _int], _int),
    "egs_problem_set_observations": ([_vp, _dp, _i64], _int),
    "egs_problem_get_observations": ([_vp, _dp, _i64], _int),
    "egs_loglik": ([_vp, _dp, _dp, _dp, _dp], _int),
    "egs_loglik_batch": ([ctypes.POINTER(_vp), _dp, _int, _dp, _dp, _dp, ctypes.POINTER(_int)], _int),
    "egs_covariance": ([_vp, _dp], _int),
    "egs_potrf": ([_vp], _int),
    "egs_solve_logdet_quad": ([_vp, _dp, _dp], _int),
    "egs_get_matrix": ([_vp, _dp, _i64], _int),
    "egs_trsm_vector": ([_vp, _int, ctypes.c_double, _dp, _i64], _int),
    "egs_last_timings": ([_vp, _dp, _dp, _dp], _int),
    "egs_last_refinement": ([_vp, ctypes.POINTER(_int), _dp], _int),
    "egs_problem_set_fallback": ([_vp, _int, ctypes.c_double], _int),
    "egs_last_fallback": ([_vp, ctypes.POINTER(_int), ctypes.POINTER(_i64), _dp], _int),
    "egs_last_product_counts": ([_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)], _int),
    "egs_last_product_counts_fp8": ([_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)], _int),
    "egs_problem_profile": ([_vp, _int], _int),
    "egs_problem_kernel_stats": ([_vp, _int, ctypes.POINTER(_i64), _dp, _dp], _int),
    "egs_generate_locations": ([_i64, _int, _dp, _dp, _dp], _int),
    "egs_generate_observations": ([_vp, _dp, _int, _dp], _int),
    "egs_predict": ([_vp, _dp, _i64, _dp, _dp, _dp, _dp, _dp, _dp], _int),
    "egs_predict_mspe": ([_vp, _dp, _i64, _dp, _dp, _dp, _dp, _dp, _dp], _int),
    "egs_bobyqa": ([_int, _dp, _dp, _dp, ctypes.c_double, _int, ctypes.c_double, _vp, _vp, _dp,
                    ctypes.POINTER(_int), ctypes.POINTER(_int)], _int),
    "egs_data_modeling": ([_vp, _dp, _dp, _dp, ctypes.c_double, _int, _dp, _dp, ctypes.POINTER(_int)], _int),
    "egs_server_create": ([_int, _i64, _int, ctypes.POINTER(_vp)], _int),
    "egs_server_destroy": ([_vp], _int),
    "egs_server_add_region": ([_vp, _int, _i64, _dp, _dp, _dp, _dp, _int, _int, ctypes.POINTER(_int)], _int),
    "egs_server_add_region_plan": ([_vp, _int, _i64, _dp, _dp, _dp, _dp, _int, _int, _int, _int, _int,
                                    ctypes.POINTER(_int)], _int),
    "egs_server_submit_loglik": ([_vp, _int, _dp, ctypes.POINTER(_i64)], _int),
    "egs_server_submit_predict": ([_vp, _int, _dp, _i64, _dp, _dp, ctypes.POINTER(_i64)], _int),
    "egs_server_wait": ([_vp, _i64, _dp, _dp, _i64, _dp], _int),
    "egs_server_stats": ([_vp, ctypes.POINTER(_i64), _dp, _dp, _dp, ctypes.POINTER(_i64)], _int),
    "egs_server_reset_stats": ([_vp], _int),
    "egs_server_batch_stats": ([_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                ctypes.POINTER(_i64), ctypes.POINTER(_i64)], _int),
}

OBJECTIVE_FN = ctypes.CFUNCTYPE(ctypes.c_double, _dp, _vp)

# Handles still open when the interpreter exits are released here, servers
# first, then problems, then contexts, while the HIP runtime is still up.
# Left to garbage collection they would be freed in no fixed order during
# interpreter teardown (a problem after its context, device calls from
# finalizers), which the exit path must never depend on.
_open_servers = weakref.WeakSet()
_open_problems = weakref.WeakSet()
_open_contexts = weakref.WeakSet()
_exiting = False


@atexit.register
def _release_open_handles():
    global _exiting
    for group in (_open_servers, _open_problems, _open_contexts):
        for obj in list(group):
            try:
                obj.close()
            except Exception:
                pass
    _exiting = True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libexageostat_amd.so is not built (run __graft_entry__.build()); "
                               "there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in EXPORTS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(code):
    if code != SUCCESS:
        msg = lib().egs_last_error().decode()
        if code == ERR_NOT_SPD:
            raise NotPositiveDefinite(code, msg)
        raise EGSError(code, msg)


def _arr(a):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _kid(kernel):
    return KERNELS[kernel] if isinstance(kernel, str) else int(kernel)


class Context:
    """One GPU: ExaGeoStatHardware's role for this library."""

    def __init__(self, device=0):
        self.handle = None
        h = _vp()
        _check(lib().egs_init(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        _open_contexts.add(self)

    def synchronize(self):
        _check(lib().egs_synchronize(self.handle))

    def close(self):
        """egs_finalize: releases the streams; problems still open on this
        context lose their device memory (closing them later is safe)."""
        if self.handle:
            h, self.handle = self.handle, None
            _open_contexts.discard(self)
            lib().egs_finalize(h)

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def covariance_tile(self, kernel, rows, cols, row_off, col_off, loc1, loc2, theta, metric=0):
        """Kernel<T>::GenerateCovarianceMatrix on the GPU; returns a (rows, cols) array."""
        out = np.zeros(rows * cols)
        x1, px1 = _arr(loc1[0])
        y1, py1 = _arr(loc1[1])
        z1, pz1 = _arr(loc1[2] if len(loc1) > 2 else None)
        x2, px2 = _arr(loc2[0])
        y2, py2 = _arr(loc2[1])
        z2, pz2 = _arr(loc2[2] if len(loc2) > 2 else None)
        th, pth = _arr(theta)
        _check(lib().egs_covariance_tile(self.handle, _kid(kernel), out.ctypes.data_as(_dp), rows, cols, row_off,
                                         col_off, px1, py1, pz1, len(x1), px2, py2, pz2, len(x2), pth, metric))
        return out.reshape(cols, rows).T.copy()


class Problem:
    """A location set with its device-resident covariance tiles (InitiateDescriptors)."""

    def __init__(self, ctx, kernel, x, y, z=None, metric=0, dts=960):
        self.ctx = ctx
        self.kernel = _kid(kernel)
        self.p = 2 if self.kernel == KERNEL_BIVARIATE else 1
        self.x, px = _arr(x)
        self.y, py = _arr(y)
        self.z, pz = _arr(z)
        self.n_loc = len(self.x)
        self.n = self.n_loc * self.p
        self.dts = dts
        self.handle = None
        h = _vp()
        _check(lib().egs_problem_create(ctx.handle, self.kernel, self.n_loc, px, py, pz, metric, dts,
                                        ctypes.byref(h)))
        self.handle = h
        _open_problems.add(self)

    def close(self):
        if self.handle:
            h, self.handle = self.handle, None
            _open_problems.discard(self)
            lib().egs_problem_destroy(h)

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass

    def device_bytes(self):
        b = _i64()
        _check(lib().egs_problem_device_bytes(self.handle, ctypes.byref(b)))
        return b.value

    def set_precision(self, mode, band_fp64=4, band_fp32=4):
        _check(lib().egs_problem_set_precision(self.handle, mode, band_fp64, band_fp32))

    def set_single_product(self, band, tau=2.0 ** -16):
        """Blocks of tiles beyond `band` take one FP16 product instead of the
        split three when their panel rows are small (row-scale product <= tau
        times the panel's largest squared); band None or negative: never."""
        _check(lib().egs_problem_set_single_product(self.handle, -1 if band is None else int(band), float(tau)))

    def set_fp8_product(self, tau):
        """Blocks of FP32 tiles beyond the single-product band take one FP8
        (e4m3) product where the row-scale product is <= tau times the
        panel's largest squared; 0 turns it off."""
        _check(lib().egs_problem_set_fp8_product(self.handle, float(tau)))

    def set_observations(self, z):
        z, pz = _arr(z)
        _check(lib().egs_problem_set_observations(self.handle, pz, len(z)))

    def get_observations(self):
        out = np.empty(self.n)
        _check(lib().egs_problem_get_observations(self.handle, out.ctypes.data_as(_dp), self.n))
        return out

    def loglik(self, theta):
        th, pth = _arr(theta)
        llh, ld, q = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().egs_loglik(self.handle, pth, ctypes.byref(llh), ctypes.byref(ld), ctypes.byref(q)))
        return llh.value, ld.value, q.value

    def covariance(self, theta):
        th, pth = _arr(theta)
        _check(lib().egs_covariance(self.handle, pth))

    def potrf(self):
        _check(lib().egs_potrf(self.handle))

    def solve_logdet_quad(self):
        ld, q = ctypes.c_double(), ctypes.c_double()
        _check(lib().egs_solve_logdet_quad(self.handle, ctypes.byref(ld), ctypes.byref(q)))
        return ld.value, q.value

    def trsm(self, z, trans=False, alpha=1.0):
        """alpha L^{-1} z (or alpha L^{-T} z) with the current FP64 factor."""
        z = np.array(z, dtype=np.float64)
        _check(lib().egs_trsm_vector(self.handle, 1 if trans else 0, float(alpha), z.ctypes.data_as(_dp), len(z)))
        return z

    def get_matrix(self):
        """Lower tiles of the current device matrix as an (n, n) array (upper part zero)."""
        out = np.zeros(self.n * self.n)
        _check(lib().egs_get_matrix(self.handle, out.ctypes.data_as(_dp), self.n))
        return out.reshape(self.n, self.n).T.copy()

    def timings(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().egs_last_timings(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"covariance_ms": a.value, "factor_ms": b.value, "solve_ms": c.value}

    def refinement(self):
        it, res = _int(), ctypes.c_double()
        _check(lib().egs_last_refinement(self.handle, ctypes.byref(it), ctypes.byref(res)))
        return {"iterations": it.value, "residual": res.value}

    def set_fallback(self, max_iterations=4, tolerance=1e-5):
        """Mixed plans: repeat an evaluation under FP64 when its estimated relative
        llh error exceeds tolerance, its refinement needs more than max_iterations
        iterations, or the mixed factor is not SPD (0 turns a test off)."""
        _check(lib().egs_problem_set_fallback(self.handle, int(max_iterations), float(tolerance)))

    def fallback(self):
        """(the last evaluation fell back to FP64, evaluations that did,
        error estimate of the last mixed evaluation)."""
        f, t, e = _int(), _i64(), ctypes.c_double()
        _check(lib().egs_last_fallback(self.handle, ctypes.byref(f), ctypes.byref(t), ctypes.byref(e)))
        return bool(f.value), t.value, e.value

    def set_graph(self, enable=True):
        """Replay FP64-plan evaluations from a captured hipGraph."""
        _check(lib().egs_problem_set_graph(self.handle, 1 if enable else 0))

    def set_fp64_product_band(self, band):
        """Tiles of the FP64 band beyond `band` take split FP16 products into
        their FP64 storage (None or negative: FP64 products throughout)."""
        _check(lib().egs_problem_set_fp64_product_band(self.handle, -1 if band is None else int(band)))

    def product_counts(self):
        """Block updates of the last factorization with the split FP16, the
        single FP16 and the single FP8 product."""
        sp, si, s8 = _i64(), _i64(), _i64()
        _check(lib().egs_last_product_counts_fp8(self.handle, ctypes.byref(sp), ctypes.byref(si), ctypes.byref(s8)))
        return {"split": sp.value, "single": si.value, "fp8": s8.value}

    def profile(self, enable=True):
        _check(lib().egs_problem_profile(self.handle, 1 if enable else 0))

    def kernel_stats(self, kernel):
        n, ms, fl = _i64(), ctypes.c_double(), ctypes.c_double()
        _check(lib().egs_problem_kernel_stats(self.handle, kernel, ctypes.byref(n), ctypes.byref(ms),
                                              ctypes.byref(fl)))
        return {"launches": n.value, "ms": ms.value, "flops": fl.value}

    def generate_observations(self, theta, seed=0):
        th, pth = _arr(theta)
        out = np.empty(self.n)
        _check(lib().egs_generate_observations(self.handle, pth, seed, out.ctypes.data_as(_dp)))
        return out

    def predict(self, theta, xm, ym, zm=None, z_actual=None):
        th, pth = _arr(theta)
        xm, pxm = _arr(xm)
        ym, pym = _arr(ym)
        zm, pzm = _arr(zm)
        za, pza = _arr(z_actual)
        out = np.empty(len(xm) * self.p)
        mspe = ctypes.c_double()
        _check(lib().egs_predict(self.handle, pth, len(xm), pxm, pym, pzm, out.ctypes.data_as(_dp), pza,
                                 ctypes.byref(mspe)))
        return out, mspe.value

    def predict_mspe(self, theta, xm, ym, zm=None, z_actual=None):
        """Prediction with the reference's three errors: (zmiss, [mspe, mspe1, mspe2])."""
        th, pth = _arr(theta)
        xm, pxm = _arr(xm)
        ym, pym = _arr(ym)
        zm, pzm = _arr(zm)
        za, pza = _arr(z_actual)
        out = np.empty(len(xm) * self.p)
        m3 = np.empty(3)
        _check(lib().egs_predict_mspe(self.handle, pth, len(xm), pxm, pym, pzm, out.ctypes.data_as(_dp), pza,
                                      m3.ctypes.data_as(_dp)))
        return out, m3

    def data_modeling(self, lb, ub, start=None, tolerance=1e-9, max_iterations=1000):
        """Maximum-likelihood fit (StarpuRuntimeSolver<T>::ModelingOperations): BOBYQA over
        [lb, ub] from start (default lb).  Returns (theta_hat, best loglik, evaluations)."""
        lb, plb = _arr(lb)
        ub, pub = _arr(ub)
        st, pst = _arr(start)
        theta = np.empty(len(lb))
        llh, nev = ctypes.c_double(), _int()
        _check(lib().egs_data_modeling(self.handle, plb, pub, pst, tolerance, max_iterations,
                                       theta.ctypes.data_as(_dp), ctypes.byref(llh), ctypes.byref(nev)))
        return theta, llh.value, nev.value


def bobyqa(fn, x0, lb, ub, ftol_abs=0.0, maxeval=1000, rho_end=0.0):
    """Minimise fn over [lb, ub] with the library's BOBYQA (host code, no GPU needed).
    Returns (x_best, f_best, evaluations, reason) with reason 1 radius, 2 ftol, 3 maxeval."""
    x = np.array(x0, dtype=np.float64)
    lb, plb = _arr(lb)
    ub, pub = _arr(ub)
    n = len(x)
    err = []

    def _cb(px, _user):
        try:
            return float(fn(np.ctypeslib.as_array(px, shape=(n,)).copy()))
        except Exception as e:  # surfaced after the call
            err.append(e)
            return float("nan")

    cb = OBJECTIVE_FN(_cb)
    fmin, nev, reason = ctypes.c_double(), _int(), _int()
    code = lib().egs_bobyqa(n, x.ctypes.data_as(_dp), plb, pub, ftol_abs, maxeval, rho_end,
                            ctypes.cast(cb, _vp), None, ctypes.byref(fmin), ctypes.byref(nev), ctypes.byref(reason))
    if err:
        raise err[0]
    _check(code)
    return x, fmin.value, nev.value, reason.value


class Server:
    """Request loop over many regions on one GPU (egs_server_*)."""

    def __init__(self, device=0, hbm_budget_bytes=64 << 30, workers=4):
        self.handle = None
        h = _vp()
        _check(lib().egs_server_create(device, int(hbm_budget_bytes), workers, ctypes.byref(h)))
        self.handle = h
        self._p = {}
        _open_servers.add(self)

    def close(self):
        if self.handle:
            h, self.handle = self.handle, None
            _open_servers.discard(self)
            lib().egs_server_destroy(h)

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass

    def add_region(self, kernel, x, y, obs, z=None, metric=0, dts=320, mode=PRECISION_FP64, band_fp64=4,
                   band_fp32=4):
        x, px = _arr(x)
        y, py = _arr(y)
        z, pz = _arr(z)
        obs, po = _arr(obs)
        rid = _int()
        _check(lib().egs_server_add_region_plan(self.handle, _kid(kernel), len(x), px, py, pz, po, metric, dts, mode,
                                                band_fp64, band_fp32, ctypes.byref(rid)))
        self._p[rid.value] = 2 if _kid(kernel) == KERNEL_BIVARIATE else 1
        return rid.value

    def submit_loglik(self, region, theta):
        th, pth = _arr(theta)
        t = _i64()
        _check(lib().egs_server_submit_loglik(self.handle, region, pth, ctypes.byref(t)))
        return (t.value, None)

    def submit_predict(self, region, theta, xm, ym):
        th, pth = _arr(theta)
        xm, pxm = _arr(xm)
        ym, pym = _arr(ym)
        t = _i64()
        _check(lib().egs_server_submit_predict(self.handle, region, pth, len(xm), pxm, pym, ctypes.byref(t)))
        return (t.value, len(xm) * self._p[region])

    def wait(self, ticket):
        """Returns (value, zmiss or None, latency_ms) of a submitted request."""
        t, nz = ticket
        v, lat = ctypes.c_double(), ctypes.c_double()
        out = np.empty(nz) if nz else None
        _check(lib().egs_server_wait(self.handle, t, ctypes.byref(v), None if out is None else out.ctypes.data_as(_dp),
                                     nz or 0, ctypes.byref(lat)))
        return v.value, out, lat.value

    def stats(self):
        c, p50, p99, ps, rb = _i64(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), _i64()
        _check(lib().egs_server_stats(self.handle, ctypes.byref(c), ctypes.byref(p50), ctypes.byref(p99),
                                      ctypes.byref(ps), ctypes.byref(rb)))
        return {"completed": c.value, "p50_ms": p50.value, "p99_ms": p99.value, "per_second": ps.value,
                "resident_bytes": rb.value}

    def reset_stats(self):
        _check(lib().egs_server_reset_stats(self.handle))

    def batch_stats(self):
        b, br, p, ph, fr = _i64(), _i64(), _i64(), _i64(), _i64()
        _check(lib().egs_server_batch_stats(self.handle, ctypes.byref(b), ctypes.byref(br), ctypes.byref(p),
                                            ctypes.byref(ph), ctypes.byref(fr)))
        return {"batches": b.value, "batched_requests": br.value, "predicts": p.value, "predict_hits": ph.value,
                "regions_switched_to_fp64": fr.value}


def loglik_batch(problems, thetas):
    """egs_loglik_batch: [(loglik, logdet, quad) or the exception of that problem] per problem."""
    n = len(problems)
    hs = (_vp * n)(*[p.handle for p in problems])
    th = np.zeros(6 * n)
    for b, t in enumerate(thetas):
        th[6 * b:6 * b + len(t)] = t
    llh, ld, q = np.zeros(n), np.zeros(n), np.zeros(n)
    st = (_int * n)()
    _check(lib().egs_loglik_batch(hs, th.ctypes.data_as(_dp), n, llh.ctypes.data_as(_dp), ld.ctypes.data_as(_dp),
                                  q.ctypes.data_as(_dp), st))
    out = []
    for b in range(n):
        if st[b] == SUCCESS:
            out.append((llh[b], ld[b], q[b]))
        else:
            out.append((NotPositiveDefinite if st[b] == ERR_NOT_SPD else EGSError)(st[b], "problem %d failed" % b))
    return out


def estimate_bytes(kernel, n_loc, dts, mode=PRECISION_FP64, band_fp64=4, band_fp32=4, dim=2):
    """Device bytes a problem of this shape and plan holds once evaluated."""
    b = _i64()
    _check(lib().egs_problem_estimate_bytes(_kid(kernel), int(n_loc), dim, dts, mode, band_fp64, band_fp32,
                                            ctypes.byref(b)))
    return b.value


def memory_cache(limit_bytes=-1):
    """Sets the device block cache limit (negative: unchanged); returns (cached, live) bytes."""
    c, l = _i64(), _i64()
    _check(lib().egs_memory_cache(int(limit_bytes), ctypes.byref(c), ctypes.byref(l)))
    return c.value, l.value


def srand(seed):
    """Seeds the C library's rand(), as ExaGeoStat<T>::ExaGeoStatLoadData does (ExaGeoStat.cpp:76)."""
    libc = ctypes.CDLL(ctypes.util.find_library("c") or "libc.so.6")
    libc.srand(ctypes.c_uint(seed))


def generate_locations(n, dim=2, seed=None):
    """LocationGenerator<T>::GenerateLocations through the library (rand() after srand(seed))."""
    if seed is not None:
        srand(seed)
    x = np.empty(n)
    y = np.empty(n)
    z = np.empty(n) if dim == 3 else None
    _check(lib().egs_generate_locations(n, dim, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                                        None if z is None else z.ctypes.data_as(_dp)))
    return (x, y) if dim != 3 else (x, y, z)
