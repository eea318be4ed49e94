"""egs_loglik_batch: several FP64 likelihood evaluations in lockstep, one
launch per kernel and factorization step for all problems.

Each batched result equals the same problem evaluated alone with egs_loglik
(same tile size: the same kernels in the same order, so the values agree to
the last bits) and the oracle at 1e-10 relative; a rejected theta or a matrix
that is not positive definite fails only its own problem; the batch leaves
each factor in place for a following prediction.
"""
import sys

import numpy as np
import pytest

import exageostatcpp_amd as egs
from exageostatcpp_amd import capi
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


def _data(sizes, dts, seed0=40):
    rng = np.random.default_rng(seed0)
    out = []
    for r, n in enumerate(sizes):
        x, y = orc.generate_locations(n, seed=seed0 + r)
        th = [rng.uniform(0.5, 2.0), rng.uniform(0.02, 0.3), rng.uniform(0.3, 2.5)]
        z = orc.generate_observations(0, (x, y), th, dts, seed=r)
        out.append((x, y, z, th))
    return out


@pytest.mark.parametrize("dts", [320, 192])
def test_batch_equals_alone_and_oracle(ctx, dts):
    sizes = [1000, 70, 640, 1500, 333, 1000]
    data = _data(sizes, dts)
    log("batch of %d problems, dts %d" % (len(sizes), dts))
    probs = []
    for x, y, z, th in data:
        p = egs.Problem(ctx, 0, x, y, dts=dts)
        p.set_observations(z)
        probs.append(p)
    got = capi.loglik_batch(probs, [d[3] for d in data])
    for (x, y, z, th), p, g in zip(data, probs, got):
        assert not isinstance(g, Exception), g
        alone = p.loglik(th)
        ref = orc.loglik(0, (x, y), th, z, dts)
        # same kernels in the same order when ts is not a multiple of 192;
        # the 192 x 192 update of the alone path otherwise
        tol = 0.0 if dts == 320 else 1e-12
        for a, b in zip(g, alone):
            assert abs(a - b) <= tol * abs(b), (g, alone)
        assert abs(g[0] - ref["llh"]) <= 1e-10 * abs(ref["llh"])
        assert abs(g[1] - ref["logdet"]) <= 1e-10 * max(1.0, abs(ref["logdet"]))
    for p in probs:
        p.close()


def test_batch_failures_stay_with_their_problem(ctx):
    data = _data([500, 800, 300], 320, seed0=60)
    probs = []
    for x, y, z, th in data:
        p = egs.Problem(ctx, 0, x, y, dts=320)
        p.set_observations(z)
        probs.append(p)
    thetas = [data[0][3], [1.0, -0.1, 0.5], [-1.0, 0.1, 0.5]]   # invalid range; negative variance: not SPD
    got = capi.loglik_batch(probs, thetas)
    ref = orc.loglik(0, data[0][:2], data[0][3], data[0][2], 320)["llh"]
    assert abs(got[0][0] - ref) <= 1e-10 * abs(ref)
    assert isinstance(got[1], capi.EGSError) and got[1].code == capi.ERR_INVALID_ARG
    assert isinstance(got[2], capi.NotPositiveDefinite)
    for p in probs:
        p.close()


def test_prediction_after_a_batch_reuses_the_factor(ctx):
    data = _data([900, 400], 320, seed0=80)
    probs = []
    for x, y, z, th in data:
        p = egs.Problem(ctx, 0, x, y, dts=320)
        p.set_observations(z)
        probs.append(p)
    capi.loglik_batch(probs, [d[3] for d in data])
    rng = np.random.default_rng(1)
    xm, ym = rng.random(20), rng.random(20)
    for (x, y, z, th), p in zip(data, probs):
        zm, _ = p.predict(th, xm, ym)
        zo, _ = orc.krige(0, (x, y), (xm, ym), th, z, 320)
        assert np.max(np.abs(zm - zo)) < 1e-8
        p.close()


def test_batch_rejects_mixed_shapes(ctx):
    x, y = orc.generate_locations(300, seed=1)
    z = np.ones(300)
    a = egs.Problem(ctx, 0, x, y, dts=320)
    b = egs.Problem(ctx, 0, x, y, dts=192)
    a.set_observations(z)
    b.set_observations(z)
    with pytest.raises(capi.EGSError):
        capi.loglik_batch([a, b], [[1, 0.1, 0.5]] * 2)
    c = egs.Problem(ctx, 0, x, y, dts=192)
    c.set_observations(z)
    c.set_precision(egs.PRECISION_MIXED)
    with pytest.raises(capi.EGSError):
        capi.loglik_batch([b, c], [[1, 0.1, 0.5]] * 2)
    for p in (a, b, c):
        p.close()
