"""Request server (egs_server_*): many regions of uneven size on one GPU.

Every served result must equal the same request evaluated alone on its own
problem (same data, same tile size, same plan): the server adds concurrency,
batching and caching, not arithmetic.  Served results are also checked against
the oracle (FP64 plan 1e-10 relative on the log-likelihood and 1e-8 on
predictions; mixed plan 1e-6 relative).  The HBM budget forces evictions in
one test.
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


def _alone(regions, reqs, mode=egs.PRECISION_FP64, dts=DTS):
    """Each request on a fresh problem of its region, one at a time, with the
    server's plan policy: a mixed region whose tiles all lie in the FP64 band
    (4) runs FP64, and a mixed region keeps FP64 after an evaluation of it fell
    back to FP64."""
    out = []
    switched = set()
    ts = (dts + 63) // 64 * 64
    with egs.Context(0) as ctx:
        for kind, r, th, xm, ym in reqs:
            x, y, z, _ = regions[r]
            p = egs.Problem(ctx, 0, x, y, dts=dts)
            p.set_observations(z)
            nt = (len(x) + ts - 1) // ts
            mixed = mode != egs.PRECISION_FP64 and nt - 1 > 4 and r not in switched
            if mixed:
                p.set_precision(mode, 4, 4)
            if kind == "loglik":
                out.append(p.loglik(th)[0])
                if mixed and p.fallback()[0]:
                    switched.add(r)
            else:
                out.append(p.predict(th, xm, ym)[0])
            p.close()
    return out


def _served(regions, reqs, budget, workers, mode=egs.PRECISION_FP64, dts=DTS, stats=None):
    srv = capi.Server(0, budget, workers)
    ids = [srv.add_region(0, x, y, z, dts=dts, mode=mode) for x, y, z, _ in regions]
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
    if stats is not None:
        stats.update(srv.batch_stats())
    srv.close()
    return res, st


def _compare(got, want):
    for g, w in zip(got, want):
        if np.isscalar(w):
            assert g == w or abs(g - w) <= 1e-13 * abs(w), (g, w)
        else:
            assert np.max(np.abs(g - w)) <= 1e-12 * max(1.0, np.max(np.abs(w))), np.max(np.abs(g - w))


def test_served_results_equal_requests_run_alone(regions):
    """All requests are queued at once, so workers take batches of the small
    FP64 likelihood requests: batched results must equal the alone ones."""
    reqs = _requests(regions, 40, 1)
    log("serve: 40 requests over %d regions, 3 workers" % len(regions))
    bs = {}
    got, st = _served(regions, reqs, 64 << 30, 3, stats=bs)
    want = _alone(regions, reqs)
    _compare(got, want)
    assert st["completed"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0.0
    log("serve: %s" % bs)
    assert bs["batches"] > 0 and bs["batched_requests"] >= 2 * bs["batches"]
    assert bs["predicts"] == sum(1 for q in reqs if q[0] == "predict")


def test_served_mixed_plan_equals_alone_and_the_oracle(regions):
    reqs = _requests(regions, 24, 3)
    log("serve: 24 requests, mixed plan, 2 workers")
    got, _ = _served(regions, reqs, 64 << 30, 2, mode=egs.PRECISION_MIXED, dts=192)
    want = _alone(regions, reqs, mode=egs.PRECISION_MIXED, dts=192)
    _compare(got, want)
    for (kind, r, th, xm, ym), g in zip(reqs, got):
        x, y, z, _ = regions[r]
        if kind == "loglik" and len(x) <= 1200:
            ref = orc.loglik(0, (x, y), th, z, 192)["llh"]
            assert abs(g - ref) <= 1e-6 * abs(ref), (g, ref)


def test_served_fp64_results_match_the_oracle(regions):
    reqs = [q for q in _requests(regions, 30, 4) if len(regions[q[1]][0]) <= 1200][:12]
    log("serve: %d requests against the oracle" % len(reqs))
    got, _ = _served(regions, reqs, 64 << 30, 2)
    for (kind, r, th, xm, ym), g in zip(reqs, got):
        x, y, z, _ = regions[r]
        if kind == "loglik":
            ref = orc.loglik(0, (x, y), th, z, DTS)["llh"]
            assert abs(g - ref) <= 1e-10 * abs(ref), (g, ref)
        else:
            zo, _ = orc.krige(0, (x, y), (xm, ym), th, z, DTS)
            assert np.max(np.abs(g - zo)) < 1e-8


def test_small_budget_evicts_and_still_matches(regions):
    reqs = _requests(regions, 30, 2)
    fp = {n: capi.estimate_bytes(0, n, DTS) for n in SIZES}
    # room for the largest region plus a quarter: the regions together need
    # about one and a half times that
    budget = int(1.25 * max(fp.values()))
    assert sum(fp.values()) > 1.4 * budget
    log("serve: 30 requests, budget %d MB, 2 workers" % (budget >> 20))
    got, st = _served(regions, reqs, budget, 2)
    want = _alone(regions, reqs)
    _compare(got, want)
    assert st["resident_bytes"] <= budget


def test_region_larger_than_budget_is_an_error(regions):
    srv = capi.Server(0, 1 << 20, 1)
    x, y, z, th = regions[3]
    with pytest.raises(capi.EGSError) as e:
        srv.add_region(0, x, y, z, dts=DTS)
    assert e.value.code == capi.ERR_OUT_OF_MEMORY
    srv.close()


@pytest.mark.parametrize("mode,dts", [(egs.PRECISION_FP64, 320), (egs.PRECISION_MIXED, 192)])
def test_footprint_estimate_equals_device_bytes(regions, mode, dts):
    """The admission footprint is what an evaluated problem holds."""
    x, y, z, th = regions[3]
    with egs.Context(0) as ctx:
        p = egs.Problem(ctx, 0, x, y, dts=dts)
        p.set_observations(z)
        if mode != egs.PRECISION_FP64:
            p.set_precision(mode, 4, 4)
        p.loglik(th)
        held = p.device_bytes()
        p.close()
    assert held == capi.estimate_bytes(0, len(x), dts, mode, 4, 4)
