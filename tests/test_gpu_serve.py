"""Request server (egs_server_*): many regions of uneven size on one GPU.

Every served result must equal the same request evaluated alone on its own
problem (same data, same tile size): the server adds concurrency and caching,
not arithmetic.  The HBM budget forces evictions in the second test.
"""
import sys

import numpy as np
import pytest

import exageostatcpp_amd as egs
from exageostatcpp_amd import capi
from oracle import orc

pytestmark = pytest.mark.gpu

SIZES = [300, 700, 1200, 2000, 450]
DTS = 320


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def regions():
    out = []
    rng = np.random.default_rng(5)
    for r, n in enumerate(SIZES):
        x, y = orc.generate_locations(n, seed=100 + r)
        th = [rng.uniform(0.5, 2.0), rng.uniform(0.02, 0.3), rng.uniform(0.3, 2.5)]
        z = orc.generate_observations(0, (x, y), th, DTS, seed=r)
        out.append((x, y, z, th))
    return out


def _requests(regions, count, seed):
    rng = np.random.default_rng(seed)
    last = {}
    reqs = []
    for _ in range(count):
        r = int(rng.integers(len(regions)))
        if r in last and rng.random() < 0.3:
            reqs.append(("predict", r, last[r], rng.random(25), rng.random(25)))
        else:
            th = [rng.uniform(0.5, 2.0), rng.uniform(0.02, 0.3), rng.uniform(0.3, 2.5)]
            last[r] = th
            reqs.append(("loglik", r, th, None, None))
    return reqs


def _alone(regions, reqs):
    """Each request on a fresh problem of its region, one at a time."""
    out = []
    with egs.Context(0) as ctx:
        for kind, r, th, xm, ym in reqs:
            x, y, z, _ = regions[r]
            p = egs.Problem(ctx, 0, x, y, dts=DTS)
            p.set_observations(z)
            if kind == "loglik":
                out.append(p.loglik(th)[0])
            else:
                out.append(p.predict(th, xm, ym)[0])
            p.close()
    return out


def _served(regions, reqs, budget, workers):
    srv = capi.Server(0, budget, workers)
    ids = [srv.add_region(0, x, y, z, dts=DTS) for x, y, z, _ in regions]
    tickets = []
    for kind, r, th, xm, ym in reqs:
        if kind == "loglik":
            tickets.append(srv.submit_loglik(ids[r], th))
        else:
            tickets.append(srv.submit_predict(ids[r], th, xm, ym))
    res = []
    for (kind, *_), t in zip(reqs, tickets):
        v, zm, lat = srv.wait(t)
        assert lat > 0.0
        res.append(v if kind == "loglik" else zm)
    st = srv.stats()
    srv.close()
    return res, st


def _compare(got, want):
    for g, w in zip(got, want):
        if np.isscalar(w):
            assert g == w or abs(g - w) <= 1e-13 * abs(w), (g, w)
        else:
            assert np.max(np.abs(g - w)) <= 1e-12 * max(1.0, np.max(np.abs(w))), np.max(np.abs(g - w))


def test_served_results_equal_requests_run_alone(regions):
    reqs = _requests(regions, 40, 1)
    log("serve: 40 requests over %d regions, 3 workers" % len(regions))
    got, st = _served(regions, reqs, 64 << 30, 3)
    want = _alone(regions, reqs)
    _compare(got, want)
    assert st["completed"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0.0


def test_small_budget_evicts_and_still_matches(regions):
    reqs = _requests(regions, 30, 2)
    # room for the largest region plus a quarter (the server's own estimate of
    # a region's bytes): the regions together need about one and a half times that
    def est(n):
        ts = (DTS + 63) // 64 * 64
        nt = (n + ts - 1) // ts
        return nt * (nt + 1) // 2 * ts * ts * 8 + nt * ts * 64 * 8 + 8 * ts * ts + 4 * nt * ts * 8 + n * 24
    budget = int(1.25 * est(max(SIZES)))
    assert sum(est(n) for n in SIZES) > 1.4 * budget
    log("serve: 30 requests, budget %d MB, 2 workers" % (budget >> 20))
    got, st = _served(regions, reqs, budget, 2)
    want = _alone(regions, reqs)
    _compare(got, want)
    assert st["resident_bytes"] <= budget


def test_region_larger_than_budget_is_an_error(regions):
    srv = capi.Server(0, 1 << 20, 1)
    x, y, z, th = regions[3]
    rid = srv.add_region(0, x, y, z, dts=DTS)
    t = srv.submit_loglik(rid, th)
    with pytest.raises(capi.EGSError) as e:
        srv.wait(t)
    assert e.value.code == capi.ERR_OUT_OF_MEMORY
    srv.close()
