"""Mixed-precision plans (FP64 band / FP32 band / 16-bit products) against the
FP64 oracle, through the C ABI."""
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


@pytest.fixture(scope="module")
def data4k():
    # 21 tiles of 192: every band class (FP64 <= 2, FP32 3..10, 16-bit > 10) occurs
    x, y = orc.generate_locations(4000, seed=0)
    th = [1.0, 0.05, 0.5]
    z = orc.generate_observations(0, (x, y), th, 192, seed=0)
    ref = orc.loglik(0, (x, y), th, z, 192)
    return x, y, th, z, ref


def test_mixed_loglik_within_1e6(ctx, data4k):
    x, y, th, z, ref = data4k
    log("mixed (split FP16) loglik n=4000 dts=192")
    p = egs.Problem(ctx, 0, x, y, dts=192)
    p.set_observations(z)
    p.set_precision(egs.PRECISION_MIXED, 2, 10)
    p.set_fallback(0, 0.0)   # the mixed arithmetic itself, no FP64 repeat
    llh, ld, q = p.loglik(th)
    rf = p.refinement()
    rel = abs(llh - ref["llh"]) / abs(ref["llh"])
    log("  llh %.10f oracle %.10f rel %.2e logdet err %.2e quad err %.2e refine %s" %
        (llh, ref["llh"], rel, ld - ref["logdet"], q - ref["quad"], rf))
    assert rf["residual"] <= 1e-12
    assert abs(q - ref["quad"]) <= 1e-9 * abs(ref["quad"])
    assert rel <= 1e-6


@pytest.mark.parametrize("log2c", [-20, 40])
def test_mixed_plan_is_equivariant_under_power_of_two_variance(ctx, data4k, log2c):
    """Sigma(c sigma^2) = c Sigma(sigma^2) with c = 2^log2c: every step of the
    mixed plan (FP64 and FP32 storage, row-scaled FP16 parts, refinement) then
    scales exactly, so log|Sigma| shifts by n ln c and z'Sigma^{-1}z divides by
    c.  At c = 2^40 the factor's entries (~2^20) exceed the FP16 range and at
    c = 2^-20 the lo parts fall below it: a split without the row scales, or an
    epilogue that does not undo them, breaks the identity."""
    x, y, th, z, _ = data4k
    n = len(x)
    c = 2.0 ** log2c
    p = egs.Problem(ctx, 0, x, y, dts=192)
    p.set_observations(z)
    p.set_precision(egs.PRECISION_MIXED, 1, 3)
    p.set_fallback(0, 0.0)
    _, ld1, q1 = p.loglik(th)
    _, ldc, qc = p.loglik([th[0] * c, th[1], th[2]])
    log("  2^%d: logdet shift %.12f (n ln c %.12f), quad ratio %.15f" % (log2c, ldc - ld1, n * np.log(c), qc * c / q1))
    assert abs((ldc - ld1) - n * np.log(c)) <= 1e-10 * abs(n * np.log(c))
    assert abs(qc * c - q1) <= 1e-12 * abs(q1)


def test_mixed_plain_bf16_refines_quad(ctx, data4k):
    x, y, th, z, ref = data4k
    log("mixed (plain BF16) loglik n=4000 dts=192")
    p = egs.Problem(ctx, 0, x, y, dts=192)
    p.set_observations(z)
    p.set_precision(egs.PRECISION_MIXED_BF16, 2, 10)
    p.set_fallback(0, 0.0)
    llh, ld, q = p.loglik(th)
    rf = p.refinement()
    log("  llh %.10f oracle %.10f rel %.2e refine %s" % (llh, ref["llh"], abs(llh - ref["llh"]) / abs(ref["llh"]), rf))
    assert rf["residual"] <= 1e-12
    assert abs(q - ref["quad"]) <= 1e-9 * abs(ref["quad"])
    assert abs(llh - ref["llh"]) <= 1e-3 * abs(ref["llh"])


def test_mixed_fp64_only_band_equals_fp64(ctx):
    """An FP64 band that covers every tile, with FP64 products throughout, is
    the FP64 factorization; with FP64 products on the diagonal tiles only (the
    default) the other tiles keep FP64 storage but take split FP16 products."""
    x, y = orc.generate_locations(1500, seed=3)
    th = [1.0, 0.1, 1.0]
    z = orc.generate_observations(0, (x, y), th, 192, seed=1)
    ref = orc.loglik(0, (x, y), th, z, 192)
    p = egs.Problem(ctx, 0, x, y, dts=192)
    p.set_observations(z)
    p.set_precision(egs.PRECISION_MIXED, 100, 100)
    p.set_fallback(0, 0.0)
    p.set_fp64_product_band(None)
    llh, _, _ = p.loglik(th)
    assert abs(llh - ref["llh"]) <= 1e-10 * abs(ref["llh"])
    p.set_fp64_product_band(0)
    llh0, _, _ = p.loglik(th)
    log("  fp64 products everywhere %.12f, on the diagonal only %.12f, oracle %.12f" % (llh, llh0, ref["llh"]))
    assert llh0 != llh and abs(llh0 - ref["llh"]) <= 1e-6 * abs(ref["llh"])


def test_mixed_kriging_matches_oracle(ctx, data4k):
    x, y, th, z, _ = data4k
    log("mixed kriging")
    rng = np.random.default_rng(2)
    xm, ym = rng.random(40), rng.random(40)
    p = egs.Problem(ctx, 0, x, y, dts=192)
    p.set_observations(z)
    p.set_precision(egs.PRECISION_MIXED, 2, 10)
    zm, _ = p.predict(th, xm, ym)
    zo, _ = orc.krige(0, (x, y), (xm, ym), th, z, 192)
    assert np.max(np.abs(zm - zo)) < 1e-8


def test_mixed_needs_tile_multiple_of_192(ctx):
    x, y = orc.generate_locations(500, seed=1)
    p = egs.Problem(ctx, 0, x, y, dts=320)
    with pytest.raises(egs.EGSError, match="multiple of 192"):
        p.set_precision(egs.PRECISION_MIXED, 2, 10)


def _mle_fixture():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mle_n1600.json")) as fh:
        return json.load(fh)


def test_data_modeling_fp64_and_mixed_theta_hat(ctx):
    """egs_data_modeling (BOBYQA on the GPU likelihood) at n = 1600: the FP64 fit
    lands within 1e-4 of the independent optimum, the mixed fit (16-bit products
    from band 4 on) within 1e-3 of it."""
    g = _mle_fixture()
    x, y = orc.generate_locations(g["n"], seed=0)
    z = orc.generate_observations(0, (x, y), g["theta_true"], g["dts"], seed=0)
    ref = np.array(g["theta_hat"])
    log("data modeling n=1600 fp64")
    p = egs.Problem(ctx, 0, x, y, dts=g["dts"])
    p.set_observations(z)
    th, llh, nev = p.data_modeling(g["lb"], g["ub"], tolerance=1e-9, max_iterations=1000)
    log("  fp64 theta %s llh %.10f (%d evaluations)" % (th, llh, nev))
    assert np.max(np.abs(th - ref)) < 1e-4
    assert llh >= g["loglik"] - 1e-6
    p.close()
    log("data modeling n=1600 mixed")
    pm = egs.Problem(ctx, 0, x, y, dts=192)
    pm.set_observations(z)
    pm.set_precision(egs.PRECISION_MIXED, 1, 3)
    thm, llhm, nevm = pm.data_modeling(g["lb"], g["ub"], tolerance=1e-9, max_iterations=1000)
    log("  mixed theta %s llh %.10f (%d evaluations)" % (thm, llhm, nevm))
    assert np.max(np.abs(thm - ref)) < 1e-3
    assert abs(llhm - g["loglik"]) < 1e-6 * abs(g["loglik"])
    pm.close()


def test_single_product_selection_pins_both_ends(ctx, data4k):
    """The per-block choice between the split and the single FP16 product:
    tau = 0 picks the single product only for all-zero blocks, so the result
    equals the split-only plan bit for bit; tau = infinity picks it for every
    block beyond the band, so the result equals what the block counts say (all
    single) and differs from the split; the default tau = 2^-16 keeps the split
    plan's accuracy."""
    x, y, th, z, ref = data4k
    p = egs.Problem(ctx, 0, x, y, dts=192)
    p.set_observations(z)
    p.set_precision(egs.PRECISION_MIXED, 1, 3)
    p.set_fallback(0, 0.0)
    p.set_single_product(-1)
    split_only = p.loglik(th)
    assert p.product_counts() == {"split": 0, "single": 0, "fp8": 0}
    p.set_single_product(0, 0.0)
    tau0 = p.loglik(th)
    c0 = p.product_counts()
    p.set_single_product(0, float("inf"))
    tau_inf = p.loglik(th)
    cinf = p.product_counts()
    p.set_single_product(0)
    dflt = p.loglik(th)
    cd = p.product_counts()
    log("  split-only %r, tau 0 %r %s, tau inf %r %s, default %r %s" % (split_only[0], tau0[0], c0, tau_inf[0], cinf,
                                                                          dflt[0], cd))
    assert tau0 == split_only
    assert c0["split"] > 0 and c0["split"] + c0["single"] == cinf["single"] and cinf["split"] == 0
    # the default plan's smallest single-product blocks take the FP8 tier
    assert cd["split"] + cd["single"] + cd["fp8"] == cinf["single"] and cd["single"] > 0 and cd["split"] > 0
    assert tau_inf[1] != split_only[1]
    assert abs(dflt[0] - ref["llh"]) <= 1e-6 * abs(ref["llh"])
    p.close()


def test_fp8_tier_moves_single_blocks_and_keeps_accuracy(ctx, data4k):
    """The FP8 tier of the per-block update (FP32 tiles beyond the band) takes
    a subset of the single-product blocks: with tau8 = 0 none; with tau8 equal
    to the single-product tau all of them and no split block (block counts),
    and tau8 = infinity takes no more; the results stay within 1e-6 of the
    oracle both with all single-product blocks in FP8 and with the default
    tau8 = 2^-26."""
    x, y, th, z, ref = data4k
    p = egs.Problem(ctx, 0, x, y, dts=192)
    p.set_observations(z)
    p.set_precision(egs.PRECISION_MIXED, 1, 3)
    p.set_fallback(0, 0.0)
    p.set_fp8_product(0.0)
    off = p.loglik(th)
    c_off = p.product_counts()
    p.set_fp8_product(2.0 ** -16)
    eq = p.loglik(th)
    c_eq = p.product_counts()
    p.set_fp8_product(float("inf"))
    allf = p.loglik(th)
    c_all = p.product_counts()
    p.set_fp8_product(2.0 ** -26)
    dflt = p.loglik(th)
    c_d = p.product_counts()
    log("  off %r %s; tau8 = tau %r %s; all %r %s; default %r %s (oracle %r)" %
        (off[0], c_off, eq[0], c_eq, allf[0], c_all, dflt[0], c_d, ref["llh"]))
    assert c_off["fp8"] == 0
    assert c_eq["fp8"] > 0 and c_eq["split"] == c_off["split"] and c_eq["single"] + c_eq["fp8"] == c_off["single"]
    assert c_all == c_eq and allf == eq
    assert allf[1] != off[1] and abs(allf[0] - ref["llh"]) <= 1e-6 * abs(ref["llh"])
    assert c_d["split"] + c_d["single"] + c_d["fp8"] == c_off["split"] + c_off["single"]
    assert abs(dflt[0] - ref["llh"]) <= 1e-6 * abs(ref["llh"])
    p.close()


@pytest.mark.parametrize("theta", [(1.0, 0.1, 0.5), (1.5, 0.2, 1.5), (0.8, 0.3, 2.5), (2.0, 0.02, 0.3)])
def test_guarded_mixed_results_are_refined_or_fp64(ctx, theta):
    """Under the default guard a mixed evaluation either is repeated in FP64 or
    its refinement converged (relative FP64 residual <= 1e-10: a breakdown of
    the refinement is a fallback, since the error estimate then says nothing),
    and kriging agrees with the FP64 plan's to 1e-8 either way (rough to very
    smooth fields, serving-sized tiles)."""
    th = list(theta)
    n = 3000
    rng = np.random.default_rng(7)
    x, y = rng.random(n), rng.random(n)
    xm, ym = rng.random(64), rng.random(64)
    p = egs.Problem(ctx, 0, x, y, dts=384)
    p.generate_observations(th, seed=3)
    l64 = p.loglik(th)[0]
    z64, _ = p.predict(th, xm, ym)
    p.set_precision(egs.PRECISION_MIXED)
    lm = p.loglik(th)[0]
    fell, _, est = p.fallback()
    rf = p.refinement()
    log("  theta %s: fp64 %.10f mixed %.10f rel %.2e fell back %s estimate %.2e refine %s" %
        (th, l64, lm, abs(lm - l64) / abs(l64), fell, est, rf))
    if not fell:
        assert rf["residual"] <= 1e-10
    zmx, _ = p.predict(th, xm, ym)
    fell_p = p.fallback()[0]
    if not fell_p:
        assert p.refinement()["residual"] <= 1e-10
    err = np.max(np.abs(zmx - z64))
    log("  kriging max diff %.2e (fell back %s)" % (err, fell_p))
    assert err <= 1e-8 * max(1.0, np.max(np.abs(z64)))
    p.close()


def test_default_mixed_plan_at_100k_within_1e6_of_fp64(ctx):
    """The benchmark workload at full size (n = 100,000, theta = (1, 0.05, 0.5),
    dts 960): the default mixed plan's log-likelihood within 1e-6 relative of
    the FP64 plan's, its quadratic form refined to FP64 accuracy, and the
    power-of-two variance identity of the mixed plan at this size."""
    th = [1.0, 0.05, 0.5]
    x, y = egs.generate_locations(100000, seed=0)
    p = egs.Problem(ctx, 0, x, y, dts=960)
    p.generate_observations(th, seed=0)
    log("n=100000: FP64 evaluation")
    l64, d64, q64 = p.loglik(th)
    p.set_precision(egs.PRECISION_MIXED)
    log("n=100000: mixed evaluations")
    lm, dm, qm = p.loglik(th)
    fell, _, est = p.fallback()
    rel = abs(lm - l64) / abs(l64)
    log("  fp64 %.10f mixed %.10f rel %.2e quad rel %.2e, error estimate %.2e" % (l64, lm, rel, abs(qm - q64) / q64,
                                                                                 est))
    assert not fell   # the default plan's own result, not the FP64 repeat
    assert rel <= 1e-6
    assert abs(qm - q64) <= 1e-10 * q64
    p.set_fallback(0, 0.0)
    _, dc, qc = p.loglik([th[0] * 2.0 ** -10, th[1], th[2]])
    assert abs((dc - dm) - 100000 * np.log(2.0 ** -10)) <= 1e-10 * abs(100000 * np.log(2.0 ** -10))
    assert abs(qc * 2.0 ** -10 - qm) <= 1e-12 * qm
    p.close()
