"""FP64 evaluations replayed from a captured hipGraph (egs_problem_set_graph).

Each replay must equal the direct path bit for bit, with the program's own
direct work and synchronizations between the replays; a change of plan must
drop the graph and a later FP64 evaluation capture a new one.
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


def test_graph_replays_equal_direct_evaluations(ctx):
    x, y = orc.generate_locations(2000, seed=4)
    z = orc.generate_observations(0, (x, y), [1.0, 0.1, 0.5], 320, seed=2)
    thetas = [[1.0, 0.1, 0.5], [1.3, 0.07, 0.8], [0.7, 0.2, 1.5], [1.0, 0.1, 0.5]]
    direct = egs.Problem(ctx, 0, x, y, dts=320)
    direct.set_observations(z)
    want = [direct.loglik(th) for th in thetas]
    graphed = egs.Problem(ctx, 0, x, y, dts=320)
    graphed.set_observations(z)
    graphed.set_graph(True)
    for rep in range(3):
        for th, w in zip(thetas, want):
            got = graphed.loglik(th)
            assert got == w, (rep, th, got, w)
            # direct work of the program between replays
            assert direct.loglik(thetas[(rep + 1) % len(thetas)]) == want[(rep + 1) % len(thetas)]
            ctx.synchronize()
    log("graph: 12 replays equal the direct path")
    # kriging reuses the factor the replay left
    xm, ym = np.random.default_rng(3).random(20), np.random.default_rng(4).random(20)
    graphed.loglik(thetas[1])
    direct.loglik(thetas[1])
    zg, _ = graphed.predict(thetas[1], xm, ym)
    zd, _ = direct.predict(thetas[1], xm, ym)
    assert np.array_equal(zg, zd)
    direct.close()
    graphed.close()


@pytest.mark.parametrize("n", [300, 700])
def test_graph_small_problems(ctx, n):
    # one tile (no bulk stream in the sequence) and three tiles
    x, y = orc.generate_locations(n, seed=9)
    th = [1.2, 0.15, 0.9]
    z = orc.generate_observations(0, (x, y), th, 320, seed=5)
    ref = orc.loglik(0, (x, y), th, z, 320)
    p = egs.Problem(ctx, 0, x, y, dts=320)
    p.set_observations(z)
    p.set_graph(True)
    for _ in range(3):
        llh = p.loglik(th)[0]
        assert abs(llh - ref["llh"]) <= 1e-10 * abs(ref["llh"])
    p.close()


def test_graph_dropped_on_plan_change_and_recaptured(ctx):
    x, y = orc.generate_locations(1500, seed=6)
    th = [1.0, 0.1, 0.5]
    z = orc.generate_observations(0, (x, y), th, 192, seed=3)
    ref = orc.loglik(0, (x, y), th, z, 192)
    p = egs.Problem(ctx, 0, x, y, dts=192)
    p.set_observations(z)
    p.set_graph(True)
    a = p.loglik(th)[0]
    p.set_precision(egs.PRECISION_MIXED)
    m = p.loglik(th)[0]
    p.set_precision(egs.PRECISION_FP64)
    b = p.loglik(th)[0]
    c = p.loglik(th)[0]
    log("graph: fp64 %.12f, mixed %.12f, fp64 again %.12f %.12f, oracle %.12f" % (a, m, b, c, ref["llh"]))
    assert a == b == c
    assert abs(a - ref["llh"]) <= 1e-10 * abs(ref["llh"])
    assert abs(m - ref["llh"]) <= 1e-6 * abs(ref["llh"])
    p.close()


def test_not_spd_reported_through_the_graph(ctx):
    # a negative variance makes Sigma negative definite: the replayed
    # factorization reports it as the direct path does, and the graph stays usable
    x, y = orc.generate_locations(400, seed=1)
    z = np.random.default_rng(0).standard_normal(400)
    p = egs.Problem(ctx, 0, x, y, dts=320)
    p.set_observations(z)
    p.set_graph(True)
    good = p.loglik([1.0, 0.1, 0.5])
    with pytest.raises(egs.NotPositiveDefinite, match="not positive definite"):
        p.loglik([-1.0, 0.1, 0.5])
    assert p.loglik([1.0, 0.1, 0.5]) == good
    p.close()


def test_graph_replay_leaves_no_covariance_to_factor(ctx):
    """After a replayed evaluation the tiles hold the factor: egs_potrf refuses
    them (it factors only the matrix of the last egs_covariance call), even
    when egs_covariance ran before the replay."""
    x, y = orc.generate_locations(500, seed=6)
    th = [1.0, 0.1, 0.5]
    p = egs.Problem(ctx, 0, x, y, dts=128)
    p.set_observations(np.linspace(-1.0, 1.0, 500))
    p.set_graph(True)
    p.loglik(th)               # captures
    p.covariance(th)
    p.loglik(th)               # replays over the generated covariance
    with pytest.raises(egs.EGSError, match="call egs_covariance first"):
        p.potrf()
    p.covariance(th)
    p.potrf()
    p.close()
