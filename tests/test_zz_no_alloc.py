"""egs_covariance_tile_noalloc: the plugin tile call without any device
allocation (a static device scratch array) gives the same tile as
egs_covariance_tile, on two calls in a row, the second smaller, and refuses a
tile whose staging exceeds the scratch."""
import ctypes

import numpy as np
import pytest

from exageostatcpp_amd import capi

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _tile(fn, ctx, kernel, rows, cols, r0, c0, loc1, loc2, theta, metric=0):
    out = np.full(rows * cols, np.nan)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (*loc1, *loc2, theta)]
    p = [v.ctypes.data_as(_dp) for v in a]
    code = fn(ctx.handle, kernel, out.ctypes.data_as(_dp), rows, cols, r0, c0, p[0], p[1], None, len(a[0]), p[2], p[3],
              None, len(a[2]), p[4], metric)
    return code, out


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("kernel,theta", [(0, [1.0, 0.1, 0.5]), (1, [1.0, 1.0, 0.1, 0.5, 1.0, 0.5])])
def test_noalloc_tile_equals_allocating_tile(ctx, kernel, theta):
    L = capi.lib()
    rng = np.random.default_rng(3)
    x, y = rng.random(3000), rng.random(3000)
    loc = (x, y)
    # a full 960 x 960 tile away from the origin, then a smaller ragged one
    for rows, cols, r0, c0 in ((960, 960, 1920, 960), (130, 66, 64, 2)):
        print("noalloc tile kernel=%d rows=%d cols=%d" % (kernel, rows, cols), flush=True)
        c1, a = _tile(L.egs_covariance_tile, ctx, kernel, rows, cols, r0, c0, loc, loc, theta)
        c2, b = _tile(L.egs_covariance_tile_noalloc, ctx, kernel, rows, cols, r0, c0, loc, loc, theta)
        assert c1 == 0 and c2 == 0, L.egs_last_error()
        assert np.array_equal(a, b)


def test_noalloc_refuses_a_tile_beyond_the_scratch(ctx):
    L = capi.lib()
    rng = np.random.default_rng(4)
    x, y = rng.random(6000), rng.random(6000)
    code, _ = _tile(L.egs_covariance_tile_noalloc, ctx, 0, 2100, 2100, 0, 0, (x, y), (x, y), [1.0, 0.1, 0.5])
    assert code == capi.ERR_INVALID_ARG
    assert b"static scratch" in L.egs_last_error()
    # the next call in range still works
    code, out = _tile(L.egs_covariance_tile_noalloc, ctx, 0, 8, 8, 0, 0, (x, y), (x, y), [1.0, 0.1, 0.5])
    assert code == 0 and np.all(np.isfinite(out)) and np.allclose(np.diag(out.reshape(8, 8)), 1.0)
