"""Benchmark: Gaussian log-likelihood evaluations per second on one MI355X.

Workload (one step = one full likelihood evaluation, ExaGeoStatMLETile):
univariate stationary Matern, theta = (1.0, 0.05, 0.5), n = 100,000 locations
on a jittered 317 x 317 grid in [0,1]^2 (glibc rand() after srand(0), Morton
order), observations z = L(theta) e generated on the GPU in FP64 (dlarnv
stream, seed 0), dense tile size 960.  Each step regenerates Sigma(theta) in
HBM, factors it, and reduces log|Sigma| and z' Sigma^-1 z; inputs are resident
in HBM when the timed region starts.

The headline value is the library's default mixed-precision plan: tiles within
band 4 of the diagonal stored in FP64 (FP64 products on the diagonal tiles,
row-scaled split-FP16 products on the others), the rest stored in FP32 and
updated per 192 x 192 block with the split-FP16 product or, where the panel
rows are small, one FP16 product, accumulated in FP32; z' Sigma^-1 z refined
to an FP64 residual <= 1e-12.  The same data is first
timed with the FP64 plan; the line reports both rates, the speedup and the
relative log-likelihood error of the mixed plan against FP64.

The roofline entry is the dominant kernel of the mixed run (the update class
with the largest time): algorithmic flops per launch over its average launch
time, timed with HIP events on the stream it runs on.  The CPU baseline times
FP64 evaluations on the host cores (the oracle's covariance tiles, host LAPACK
dpotrf and dtrtrs) on a bounded sample and scales the median to n by the n^3
(factorization) and n^2 (covariance, solve) laws.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "llh evals/s at n=100k, mixed vs FP64 GPU (rel llh err ≤1e-6); serve p50/p99"
PEAK = {"fp64": 78.6, "fp32": 157.3, "f16": 2500.0, "bf16": 2500.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """(threads this process may use, physical cores of the host from lscpu)."""
    affinity = len(os.sched_getaffinity(0))
    physical = None
    try:
        import subprocess

        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        physical = len({l for l in out.splitlines() if l and not l.startswith("#")})
    except (OSError, ValueError):
        pass
    return affinity, physical


def cpu_baseline(n_target, n_sample, dts, theta, reps=3):
    """FP64 likelihood evaluations on the host cores: the oracle's tiled
    Matern covariance (C, OpenMP over tiles), host LAPACK dpotrf (OpenBLAS,
    multithreaded) for the factorization, then log|Sigma| from the diagonal and
    one triangular solve with dtrtrs.  One warm-up and `reps` timed
    evaluations at n_sample; the median is scaled to n_target by n^3 for the
    factorization and n^2 for covariance and solve."""
    import numpy as np
    import scipy.linalg.blas as blas
    import scipy.linalg.lapack as lapack
    import threadpoolctl

    from oracle import orc

    affinity, physical = host_cores()
    threads = min(16, affinity)
    x, y = orc.generate_locations(n_sample, seed=0)
    z = None
    runs = []
    with threadpoolctl.threadpool_limits(threads):
        for rep in range(reps + 1):
            t0 = time.perf_counter()
            C = orc.covariance_matrix(0, (x, y), (x, y), theta, dts, order="F")
            t1 = time.perf_counter()
            L, info = lapack.dpotrf(C, lower=1, clean=0, overwrite_a=1)
            t2 = time.perf_counter()
            if info != 0:
                raise RuntimeError("host dpotrf failed (info %d)" % info)
            if z is None:
                # observations z = L e from the warm-up factor (LinearAlgebraMethods.cpp:333-449)
                e = orc.dlarnv_normal([0, 0, 0, 1], n_sample)[0]
                z = blas.dtrmv(L, e, lower=1)
            t3 = time.perf_counter()
            logdet = 2.0 * np.sum(np.log(np.diag(L)))
            w, info = lapack.dtrtrs(L, z, lower=1)
            quad = float(w @ w)
            llh = -0.5 * quad - 0.5 * logdet - 0.5 * n_sample * np.log(2.0 * np.pi)
            t4 = time.perf_counter()
            del C, L
            if rep > 0:
                runs.append((t1 - t0, t2 - t1, t4 - t3, t4 - t0 - (t3 - t2)))
            log("cpu baseline rep %d: covariance %.2f s, dpotrf %.2f s, solve %.3f s, llh %.6f"
                % (rep, t1 - t0, t2 - t1, t4 - t3, llh))
    med = sorted(runs, key=lambda r: r[3])[len(runs) // 2]
    s = n_target / n_sample
    t_scaled = med[1] * s ** 3 + (med[0] + med[2]) * s ** 2
    return {
        "value": 1.0 / t_scaled,
        "unit": "llh evals/s",
        "cores": threads,
        "host_physical_cores": physical,
        "kind": "port",
        "measured_evals_per_s_at_sample": 1.0 / med[3],
        "sample": "median of %d FP64 likelihood evaluations (after one warm-up) at n=%d, same theta: oracle "
                  "covariance (C, OpenMP over %d-tiles) %.2f s, host LAPACK dpotrf (OpenBLAS, %d threads) %.2f s "
                  "(%.0f GFLOP/s), log-det and dtrtrs %.3f s; scaled to n=%d by n^3 for the factorization and n^2 "
                  "for covariance and solve (a restatement on the host cores, not Chameleon+StarPU)"
                  % (reps, n_sample, dts, med[0], threads, med[1], n_sample ** 3 / 3.0 / med[1] / 1e9, med[2],
                     n_target),
    }


def operands(args):
    """Operand format of the 16-bit update class: 'fp16-split' (row-scaled
    hi/lo FP16, three products) or 'bf16' (one product)."""
    return "bf16" if args.plain_bf16 else "fp16-split"


def fp8_tau(args):
    """The FP8-product threshold in effect for the univariate kernel: the
    option, else the library default (2^-26; 0 turns the tier off)."""
    return args.fp8_tau if args.fp8_tau is not None else 2.0 ** -26


def product_band(args, bivariate):
    """The FP64-product band in effect: the option, else the library default."""
    if args.fp64_product_band is not None:
        return args.fp64_product_band
    return 2 if bivariate else 0


def tau_json(args):
    """--single-tau for the JSON line (a string for infinity, which JSON lacks)."""
    return "inf" if args.single_tau == float("inf") else args.single_tau


def measured_traffic(kernel, args):
    """(HBM bytes per launch, algorithmic bytes per launch) of `kernel` from
    the committed rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated), when it was taken on this same workload; (None, None)
    otherwise."""
    path = os.path.join(ROOT, "profiles", "%s_n%dk_traffic.json" % (kernel, args.n // 1000))
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None, None
    same = (t.get("kernel") == kernel and t.get("n") == args.n and t.get("dts") == args.dts
            and t.get("band_fp64") == args.band_fp64 and t.get("band_fp32") == args.band_fp32
            and t.get("single_band") == args.single_band and t.get("single_tau") == tau_json(args)
            and t.get("fp64_product_band", -1) == product_band(args, bivariate=False)
            and t.get("operands") == operands(args) and t.get("fp8_tau", 0.0) == fp8_tau(args))
    if not same:
        return None, None
    return t["fetch_bytes_per_launch"] + t["write_bytes_per_launch"], t.get("algorithmic_bytes_per_launch")


def _counter_mean_kib(path, name_regex):
    """Mean Counter_Value (KiB) per dispatch of the kernels matching
    name_regex in a rocprofv3 --pmc CSV, and the number of dispatches."""
    import csv
    import re

    per = {}
    with open(path) as fh:
        for r in csv.DictReader(fh):
            if re.search(name_regex, r["Kernel_Name"]):
                per[r["Dispatch_Id"]] = per.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    if not per:
        raise SystemExit("no dispatch of %s in %s" % (name_regex, path))
    return sum(per.values()) / len(per), len(per)


def traffic_summary(args):
    """Writes profiles/update_16_select_n<N>k_traffic.json from four rocprofv3
    --pmc runs under DIR (f: FETCH_SIZE and w: WRITE_SIZE of the bench
    workload, cf and cw: the same counters of --copy-probe's 1 GiB copy), each
    in DIR/<x>/<x>_counter_collection.csv.  The copy calibrates the counters
    to bytes; the algorithmic bytes per launch are the per-block class's tile
    reads and writes (one FP32 tile read and written per tile update) over the
    schedule's launches of that class."""
    d, rx = args.traffic_summary, args.traffic_kernel
    f_kib, nf = _counter_mean_kib(os.path.join(d, "f", "f_counter_collection.csv"), rx)
    w_kib, nw = _counter_mean_kib(os.path.join(d, "w", "w_counter_collection.csv"), rx)
    cf_kib, _ = _counter_mean_kib(os.path.join(d, "cf", "cf_counter_collection.csv"), "copy")
    cw_kib, _ = _counter_mean_kib(os.path.join(d, "cw", "cw_counter_collection.csv"), "copy")
    gib = float(2 ** 30)
    cal_f, cal_w = gib / (cf_kib * 1024.0), gib / (cw_kib * 1024.0)
    # the per-block class: tiles with i - j > max(band_fp32, single_band);
    # launches per step: column k+1 (critical), column k+2, columns k+3.. (bulk)
    nt = -(-args.n // args.dts)
    beyond = max(args.band_fp32, args.single_band)
    launches, updates = 0, 0
    for k in range(nt - 1):
        for lo, hi in ((k + 1, k + 2), (k + 2, k + 3), (k + 3, nt)):
            cnt = sum(max(0, nt - j - beyond - 1) for j in range(lo, min(hi, nt)))
            if cnt:
                launches += 1
                updates += cnt
    alg = updates * 2.0 * 4.0 * args.dts * args.dts / launches
    out = {"kernel": "update_16_select", "kernel_name_regex": rx, "n": args.n, "dts": args.dts,
           "band_fp64": args.band_fp64, "band_fp32": args.band_fp32, "single_band": args.single_band,
           "single_tau": tau_json(args), "fp64_product_band": product_band(args, bivariate=False),
           "operands": operands(args), "fp8_tau": fp8_tau(args),
           "fetch_bytes_per_launch": f_kib * 1024.0 * cal_f, "write_bytes_per_launch": w_kib * 1024.0 * cal_w,
           "algorithmic_bytes_per_launch": alg, "launches_per_factorization": launches,
           "tile_updates_per_factorization": updates, "calibration": {"fetch": cal_f, "write": cal_w},
           "raw_kib_per_dispatch": {"fetch": f_kib, "write": w_kib}, "dispatches": {"fetch": nf, "write": nw}}
    path = args.traffic_out or os.path.join(ROOT, "profiles", "update_16_select_n%dk_traffic.json" % (args.n // 1000))
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def timed(egs, ctx, prob, theta, steps, warmup, label):
    """Warm-up, then `steps` timed evaluations without per-kernel timing, then
    one more evaluation with HIP-event timing of every launch (kernel table and
    roofline; the event bookkeeping costs host time, so it stays out of the
    timed steps)."""
    for i in range(warmup):
        t0 = time.perf_counter()
        prob.loglik(theta)
        log("%s warmup %d: %.3f s" % (label, i, time.perf_counter() - t0))
    ctx.synchronize()
    phases = {"covariance_ms": 0.0, "factor_ms": 0.0, "solve_ms": 0.0}
    res = None
    t0 = time.perf_counter()
    for i in range(steps):
        res = prob.loglik(theta)
        for k, v in prob.timings().items():
            phases[k] += v
        log("%s step %d: llh %.10f" % (label, i, res[0]))
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    prob.profile(True)
    prob.loglik(theta)
    ctx.synchronize()
    kinds = (("update_fp64", egs.capi.PROFILE_UPDATE), ("update_fp32", egs.capi.PROFILE_UPDATE_F32),
             ("update_16", egs.capi.PROFILE_UPDATE_16), ("update_16_select", egs.capi.PROFILE_UPDATE_16_SELECT),
             ("covariance", egs.capi.PROFILE_COVARIANCE),
             ("panel_trsm", egs.capi.PROFILE_PANEL_TRSM), ("diagonal", egs.capi.PROFILE_DIAGONAL),
             ("forward_solve", egs.capi.PROFILE_SOLVE), ("refine", egs.capi.PROFILE_REFINE))
    ks = {name: prob.kernel_stats(kid) for name, kid in kinds}
    prob.profile(False)
    return {
        "evals_per_s": steps / elapsed,
        "ms_per_step": 1000.0 * elapsed / steps,
        "loglik": res[0], "logdet": res[1], "quad": res[2],
        "phases_ms_per_step": {k: v / steps for k, v in phases.items()},
        "kernels": ks,
        "refinement": prob.refinement(),
        "product_counts": prob.product_counts(),
    }


def kernel_table(ks, steps):
    out = {}
    for k, v in ks.items():
        if v["launches"] == 0:
            continue
        out[k] = {"ms": v["ms"] / steps, "launches": v["launches"] // steps,
                  "tflops": (v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 and v["flops"] > 0 else None}
    return out


# Tile size of the serving regions: a multiple of 192, as mixed plans need, and
# the size the observations are generated with, so that an FP64 evaluation at
# the generating theta repeats the generation's factorization.
SERVE_DTS = 384


def serve_workload(egs, args):
    """The serving mix: args.serve_requests requests; 80% likelihood evaluations,
    each for a new region with n log-uniform in [500, 20000] (jittered-grid
    locations, theta uniform in [0.5, 2] x [0.02, 0.3] x [0.3, 2.5], z simulated
    at theta with the FP64 factor on the GPU), 20% kriging of 100 new uniform
    locations against one of the 16 regions most recently submitted (at its
    theta, so the factor cached on the device is reused); Poisson arrival times
    at args.serve_rate per second."""
    import numpy as np

    rng = np.random.default_rng(7)
    reqs, regions = [], []
    for i in range(args.serve_requests):
        if regions and rng.random() < 0.2:
            r = int(rng.integers(max(0, len(regions) - 16), len(regions)))
            reqs.append(("predict", r, rng.random(100), rng.random(100)))
        else:
            n = int(np.exp(rng.uniform(np.log(500), np.log(20000))))
            th = [rng.uniform(0.5, 2.0), rng.uniform(0.02, 0.3), rng.uniform(0.3, 2.5)]
            regions.append([n, th, 10000 + len(regions), None, None, None])
            reqs.append(("loglik", len(regions) - 1, None, None))
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, len(reqs)))
    t0 = time.perf_counter()
    egs.capi.memory_cache(64 << 30)
    with egs.Context(0) as ctx:
        for reg in regions:
            n, th, seed = reg[0], reg[1], reg[2]
            x, y = egs.generate_locations(n, seed=seed)
            p = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=SERVE_DTS)
            reg[3], reg[4], reg[5] = x, y, p.generate_observations(th, seed=seed)
            p.close()
    egs.capi.memory_cache(0)
    log("serve: %d regions (n %d..%d) generated in %.1f s" % (len(regions), min(r[0] for r in regions),
                                                              max(r[0] for r in regions), time.perf_counter() - t0))
    return reqs, regions, arrivals


def serve_run(egs, args, reqs, regions, arrivals, mode):
    """Runs the serving mix through one server under `mode`; returns the
    summary and the per-request results."""
    import numpy as np

    mixed = mode != egs.PRECISION_FP64
    dts = SERVE_DTS
    srv = egs.capi.Server(0, args.serve_budget_gb << 30, args.serve_workers)
    ids = [srv.add_region(0, r[3], r[4], r[5], dts=dts, mode=mode, band_fp64=args.band_fp64,
                          band_fp32=args.band_fp32) for r in regions]
    # warm-up on two extra small regions (kernel loading), then fresh statistics
    rng = np.random.default_rng(11)
    for k in range(2):
        x, y = egs.generate_locations(600, seed=90 + k)
        w = srv.add_region(0, x, y, rng.standard_normal(600), dts=dts, mode=mode)
        srv.wait(srv.submit_loglik(w, [1.0, 0.1, 0.5]))
    srv.reset_stats()
    tickets = []
    start = time.perf_counter()
    for (kind, r, xm, ym), ta in zip(reqs, arrivals):
        dt = start + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        th = regions[r][1]
        if kind == "predict":
            tickets.append(srv.submit_predict(ids[r], th, xm, ym))
        else:
            tickets.append(srv.submit_loglik(ids[r], th))
    out, failed = [], []
    for i, t in enumerate(tickets):
        try:
            v, zm, _ = srv.wait(t)
            out.append(zm if reqs[i][0] == "predict" else v)
        except egs.EGSError as e:   # e.g. a covariance that is not positive definite at that theta
            out.append(None)
            failed.append((i, str(e)))
    wall = time.perf_counter() - start
    st = srv.stats()
    bs = srv.batch_stats()
    srv.close()
    log("serve %s: %d requests at %.0f/s: p50 %.1f ms, p99 %.1f ms, %d batches holding %d requests, "
        "%d of %d predictions on a cached factor" % ("mixed" if mixed else "fp64", len(tickets), args.serve_rate,
                                                    st["p50_ms"], st["p99_ms"], bs["batches"], bs["batched_requests"],
                                                    bs["predict_hits"], bs["predicts"]))
    if failed:
        log("serve: %d requests failed, first: %s" % (len(failed), failed[0][1]))
    return {"plan": ("mixed (%d, %d) with its FP64 fallback, dts %d" % (args.band_fp64, args.band_fp32, dts)) if mixed
            else "fp64, dts %d" % dts,
            "failed": len(failed), "failure_example": failed[0][1] if failed else None,
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "completed_per_s": len(tickets) / wall,
            "batches": bs["batches"], "batched_requests": bs["batched_requests"],
            "regions_switched_to_fp64": bs["regions_switched_to_fp64"],
            "predictions_on_cached_factor": [bs["predict_hits"], bs["predicts"]]}, out


def serve_phase(egs, args):
    """Serving at a fixed request rate under the FP64 and the mixed plan (see
    serve_workload); latency is submission to completion.  The per-request
    results of the two plans are compared (relative log-likelihood error and
    largest prediction difference of mixed against FP64)."""
    import numpy as np

    reqs, regions, arrivals = serve_workload(egs, args)
    fp64, r64 = serve_run(egs, args, reqs, regions, arrivals, egs.PRECISION_FP64)
    mixed, rmx = serve_run(egs, args, reqs, regions, arrivals, egs.PRECISION_MIXED)
    both = [(k, a, b) for (k, *_), a, b in zip(reqs, rmx, r64) if a is not None and b is not None]
    llh_err = max((abs(a - b) / abs(b) for k, a, b in both if k == "loglik"), default=0.0)
    pred_err = max((float(np.max(np.abs(a - b))) for k, a, b in both if k == "predict"), default=0.0)
    sizes = [r[0] for r in regions]
    return {"requests": len(reqs), "rate_per_s": args.serve_rate, "regions": len(regions),
            "n_range": [min(sizes), max(sizes)], "workers": args.serve_workers,
            "hbm_budget_gb": args.serve_budget_gb,
            "mix": "80% likelihood of a new region each (n log-uniform in [500, 20000], theta uniform, z simulated "
                   "at theta), 20% kriging of 100 new points against one of the 16 latest regions at its theta",
            "p50_ms": fp64["p50_ms"], "p99_ms": fp64["p99_ms"], "fp64": fp64, "mixed": mixed,
            "mixed_vs_fp64": {"max_rel_llh_err": llh_err, "max_abs_prediction_diff": pred_err}}


def bivariate_gcd(egs, args):
    """The bivariate parsimonious Matérn workload: 25,000 locations uniform on
    the sphere (latitude, longitude in degrees), haversine distance in km,
    theta = (1, 1, 637.1 km, 0.5, 1.0, 0.5), dts 960 (matrix dimension 50,000);
    FP64 and mixed evaluations per second."""
    import numpy as np

    n_loc = 25000 if args.n == 100000 else args.n
    rng = np.random.default_rng(0)
    u, v = rng.random(n_loc), rng.random(n_loc)
    lat, lon = np.degrees(np.arcsin(2.0 * u - 1.0)), 360.0 * v - 180.0
    # Morton order on a 16-bit grid of (longitude, latitude), as the reference's
    # generator sorts its locations, so that tile distance follows spatial distance
    qx = np.round((lon + 180.0) / 360.0 * 65535).astype(np.uint64)
    qy = np.round((lat + 90.0) / 180.0 * 65535).astype(np.uint64)
    code = np.zeros(n_loc, dtype=np.uint64)
    for b in range(16):
        code |= ((qx >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b)
        code |= ((qy >> np.uint64(b)) & np.uint64(1)) << np.uint64(2 * b + 1)
    order = np.argsort(code, kind="stable")
    lat, lon = lat[order], lon[order]
    theta = [1.0, 1.0, 637.1, 0.5, 1.0, 0.5]
    ctx = egs.Context(0)
    prob = egs.Problem(ctx, "BivariateMaternParsimonious", lat, lon, metric=1, dts=args.dts)
    prob.generate_observations(theta, seed=0)
    out = {"workload": "bivariate parsimonious Matern, %d locations on the sphere, great-circle km, theta=%s, "
                       "dts=%d (dimension %d), Morton-ordered" % (n_loc, theta, args.dts, 2 * n_loc)}
    fp64 = timed(egs, ctx, prob, theta, args.steps, args.warmup, "bivariate fp64")
    out["fp64"] = {"evals_per_s": fp64["evals_per_s"], "ms_per_step": fp64["ms_per_step"], "loglik": fp64["loglik"],
                   "tflops": (2 * n_loc) ** 3 / 3.0 / (fp64["ms_per_step"] * 1e-3) / 1e12}
    out["fp64_product_band"] = product_band(args, bivariate=True)
    out["mixed"] = []
    for b64, b32 in ((1, 1), (2, 2), (4, 4), (8, 8)):
        prob.set_precision(egs.PRECISION_MIXED, b64, b32)
        prob.set_single_product(args.single_band, args.single_tau)
        if args.fp64_product_band is not None:
            prob.set_fp64_product_band(args.fp64_product_band)
        # the plan's own arithmetic (guard off), then one evaluation under the
        # default guard: its error estimate and whether it repeats in FP64
        prob.set_fallback(0, 0.0)
        try:
            mixed = timed(egs, ctx, prob, theta, args.steps, args.warmup, "bivariate mixed (%d, %d)" % (b64, b32))
        except egs.NotPositiveDefinite as e:
            log("bivariate mixed (%d, %d): %s" % (b64, b32, e))
            out["mixed"].append({"plan": [b64, b32], "error": str(e)})
            prob.set_fallback()
            continue
        prob.set_fallback()
        prob.loglik(theta)
        fell, _, est = prob.fallback()
        out["mixed"].append({"plan": [b64, b32], "evals_per_s": mixed["evals_per_s"],
                             "ms_per_step": mixed["ms_per_step"], "loglik": mixed["loglik"],
                             "rel_llh_err_vs_fp64": abs(mixed["loglik"] - fp64["loglik"]) / abs(fp64["loglik"]),
                             "refinement": mixed["refinement"],
                             "block_updates_beyond_single_band": mixed["product_counts"],
                             "speedup_vs_fp64": mixed["evals_per_s"] / fp64["evals_per_s"],
                             "guard": {"error_estimate": est, "repeats_in_fp64": bool(fell)}})
    prob.close()
    ctx.close()
    print(json.dumps(out))


def c2_workflow(egs, args):
    """The smooth-field workflow: n = 40,000 locations on a jittered 200 x 200
    grid (srand(0)) and 1,000 uniform held-out points with observations
    generated jointly (z = L(theta) e over all 41,000 points, theta = (1, 0.1,
    1.0), FP64 on the GPU); maximum likelihood on the 40,000 observations with
    the FP64 plan, dts 960, from the lower bounds (0.01, 0.01, 0.01) within
    [0.01, 5]^3 and at most --c2-iterations evaluations; then kriging of the
    held-out points at the estimate, with the MSPE against their generated
    values (ExaGeoStatDataModeling then ExaGeoStatPrediction)."""
    import numpy as np

    n, held, theta = 40000, 1000, [1.0, 0.1, 1.0]
    x, y = egs.generate_locations(n, seed=0)
    rng = np.random.default_rng(n + held)
    xm, ym = rng.random(held), rng.random(held)
    t0 = time.perf_counter()
    with egs.Context(0) as ctx:
        pj = egs.Problem(ctx, "UnivariateMaternStationary", np.concatenate([x, xm]), np.concatenate([y, ym]),
                         dts=args.dts)
        z = pj.generate_observations(theta, seed=0)
        pj.close()
        zo, zh = z[:n], z[n:]
        t1 = time.perf_counter()
        p = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=args.dts)
        p.set_observations(zo)
        th, llh, nev = p.data_modeling([0.01] * 3, [5.0] * 3, tolerance=args.mle_ftol,
                                       max_iterations=args.c2_iterations)
        t2 = time.perf_counter()
        zm, m3 = p.predict_mspe(list(th), xm, ym, z_actual=zh)
        t3 = time.perf_counter()
        p.close()
    log("c2: theta-hat %s llh %.10f after %d evaluations in %.1f s; MSPE %.6g over %d held-out points"
        % (list(th), llh, nev, t2 - t1, m3[0], held))
    out = {"workload": "smooth-field workflow: n=%d jittered 200 x 200 grid + %d uniform held-out points generated "
                       "jointly, theta_true=%s, FP64, dts=%d, MLE from the lower bounds in [0.01, 5]^3 with at most %d "
                       "evaluations (ftol %g), then kriging of the held-out points" % (
                           n, held, theta, args.dts, args.c2_iterations, args.mle_ftol),
           "theta_hat": list(th), "loglik": llh, "evaluations": nev, "mle_seconds": t2 - t1,
           "evals_per_s": nev / (t2 - t1), "data_seconds": t1 - t0, "mspe": m3[0], "prediction_seconds": t3 - t2,
           "held_out_variance": float(np.var(zh))}
    print(json.dumps(out))


def mle_compare(egs, args):
    """Maximum-likelihood fits of one data set with the FP64 and the mixed plan.
    Both start at the generating theta inside the box [theta/2, 2 theta] (range
    [0.4, 4] x theta) so that both converge within the evaluation budget; a point
    where the mixed factor is not positive definite is evaluated with the FP64
    plan instead (counted in the output)."""
    n = args.mle_n
    theta = args.theta or [1.0, 0.05, 0.5]
    lb = [0.5 * theta[0], 0.4 * theta[1], 0.5 * theta[2]]
    ub = [2.0 * theta[0], 4.0 * theta[1], 2.0 * theta[2]]
    x, y = egs.generate_locations(n, seed=0)
    ctx = egs.Context(0)
    prob = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=args.dts)
    prob.generate_observations(theta, seed=0)
    out = {"workload": "maximum-likelihood fit, n=%d, theta_true=%s, dts=%d, bounds %s..%s, start theta_true, "
                       "ftol %g, rho_end %g" % (n, theta, args.dts, lb, ub, args.mle_ftol, args.mle_rho_end)}
    out["z_sum"] = float(sum(prob.get_observations()))   # identical data across separate runs
    fp64_prob = None
    plans = [pl for pl in (("fp64", None), ("mixed", egs.PRECISION_MIXED)) if args.mle_plans in ("both", pl[0])]
    for label, mode in plans:
        if mode is not None:
            fp64_prob = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=args.dts)
            fp64_prob.set_observations(prob.get_observations())
            prob.set_precision(mode, args.band_fp64, args.band_fp32)
            prob.set_single_product(args.single_band, args.single_tau)
            prob.set_fp8_product(fp8_tau(args))
            if args.fp64_product_band is not None:
                prob.set_fp64_product_band(args.fp64_product_band)
        count, fallback = [0], [0]

        def objective(v):
            count[0] += 1
            try:
                llh = prob.loglik(list(v))[0]
                if mode is not None and prob.fallback()[0]:
                    fallback[0] += 1   # the library's guard repeated it under the FP64 plan
            except egs.NotPositiveDefinite:
                fallback[0] += 1
                llh = fp64_prob.loglik(list(v))[0]
            if count[0] % 10 == 0:
                log("%s evaluation %d: theta %s llh %.10f" % (label, count[0], list(v), llh))
            return -llh

        t0 = time.perf_counter()
        th, fmin, nev, _ = egs.capi.bobyqa(objective, theta, lb, ub, ftol_abs=args.mle_ftol, maxeval=args.maxeval,
                                           rho_end=args.mle_rho_end)
        dt = time.perf_counter() - t0
        out[label] = {"theta_hat": list(th), "loglik": -fmin, "evaluations": nev, "fp64_fallbacks": fallback[0],
                      "seconds": dt}
        log("%s: theta %s llh %.10f, %d evaluations, %.1f s" % (label, list(th), -fmin, nev, dt))
    if "fp64" in out and "mixed" in out:
        d = [abs(a - b) for a, b in zip(out["fp64"]["theta_hat"], out["mixed"]["theta_hat"])]
        out["max_abs_theta_diff"] = max(d)
    if "mixed" in out:
        out["fp64_loglik_at_mixed_theta_hat"] = fp64_prob.loglik(out["mixed"]["theta_hat"])[0]
        out["mixed_plan"] = {"band_fp64": args.band_fp64, "band_fp32": args.band_fp32,
                             "single_band": args.single_band, "single_tau": tau_json(args),
                             "fp64_product_band": product_band(args, bivariate=False)}
        fp64_prob.close()
    prob.close()
    ctx.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=None, help="timed FP64 steps (default: --steps)")
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--dts", type=int, default=960)
    ap.add_argument("--band-fp64", type=int, default=4)
    ap.add_argument("--band-fp32", type=int, default=4)
    ap.add_argument("--single-band", type=int, default=0,
                    help="blocks of tiles beyond this band may take one FP16 product instead of the split three "
                         "(negative: never)")
    ap.add_argument("--fp64-product-band", type=int, default=None,
                    help="FP64-band tiles beyond this band take split FP16 products into FP64 storage "
                         "(negative: every tile of the FP64 band takes FP64 products; default: the library's, 0 for "
                         "univariate and 2 for bivariate kernels)")
    ap.add_argument("--single-tau", type=float, default=2.0 ** -16,
                    help="... when their panel rows' scales multiply to <= tau times the panel's largest squared")
    ap.add_argument("--fp8-tau", type=float, default=None,
                    help="FP32-tile blocks whose panel rows' scales multiply to <= this times the panel's largest "
                         "squared take one FP8 (e4m3) product (0: never; default: the library's, 2^-26)")
    ap.add_argument("--plain-bf16", action="store_true", help="single BF16 product instead of the split one")
    ap.add_argument("--cpu-sample", type=int, default=20000,
                    help="n of the host sample (each evaluation a few seconds on 16 host cores)")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed host evaluations (median reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mle", action="store_true",
                    help="fit theta with the FP64 and the mixed plan on the same data and report both estimates "
                         "(defaults to n=40000, theta=(1, 0.05, 0.5), ftol 1e-9, tighter than the reference's "
                         "default 1e-8)")
    ap.add_argument("--mle-ftol", type=float, default=1e-9)
    ap.add_argument("--mle-rho-end", type=float, default=0.0,
                    help="BOBYQA's final trust-region radius (0: the solver's default); with --mle-ftol 0 the fits "
                         "stop on the radius alone")
    ap.add_argument("--bivariate", action="store_true",
                    help="the bivariate great-circle workload (25,000 locations on the sphere) instead of the default")
    ap.add_argument("--c2", action="store_true",
                    help="the smooth-field workflow: MLE at n=40000 (nu = 1) then kriging of 1000 jointly generated "
                         "held-out points")
    ap.add_argument("--c2-iterations", type=int, default=100)
    ap.add_argument("--mle-n", type=int, default=40000, help="n of the --mle fits")
    ap.add_argument("--mle-plans", choices=("both", "fp64", "mixed"), default="both",
                    help="which --mle fits to run (separate runs see identical data: z_sum in the output)")
    ap.add_argument("--theta", type=float, nargs=3, default=None)
    ap.add_argument("--maxeval", type=int, default=500)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving phase")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=50.0, help="request arrivals per second (Poisson)")
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-budget-gb", type=int, default=160, help="HBM budget of the server")
    ap.add_argument("--serve-only", action="store_true", help="only the serving phase")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--traffic-summary", metavar="DIR", default=None,
                    help="write the dominant kernel's HBM traffic summary under profiles/ from rocprofv3 FETCH_SIZE "
                         "and WRITE_SIZE runs in DIR/{f,w,cf,cw} (no GPU needed)")
    ap.add_argument("--traffic-kernel", default="update192_w_kernelIDF16_Li2ELi2ELb0",
                    help="kernel-name regex of --traffic-summary")
    ap.add_argument("--traffic-out", default=None, help="output file of --traffic-summary (default: under profiles/)")
    args = ap.parse_args()
    if args.traffic_summary:
        return traffic_summary(args)
    if args.copy_probe:
        import torch

        a = torch.ones(2 ** 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copy probe: 1 GiB read, 1 GiB written")
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the likelihood path is built and timed on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import exageostatcpp_amd as egs

    if args.serve_only:
        print(json.dumps(serve_phase(egs, args)))
        return
    if args.c2:
        return c2_workflow(egs, args)
    if args.mle:
        return mle_compare(egs, args)
    if args.bivariate:
        return bivariate_gcd(egs, args)

    theta = [1.0, 0.05, 0.5]
    t_setup = time.perf_counter()
    x, y = egs.generate_locations(args.n, seed=0)
    ctx = egs.Context(0)
    prob = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=args.dts)
    prob.generate_observations(theta, seed=0)
    log("setup: n=%d, data generated in %.1f s" % (args.n, time.perf_counter() - t_setup))

    fp64_steps = args.fp64_steps if args.fp64_steps is not None else args.steps
    fp64 = timed(egs, ctx, prob, theta, fp64_steps, args.warmup, "fp64")
    # kriging of 1000 uniform points with each plan's factor at theta (reused
    # from the last evaluation): the prediction error of mixed against FP64
    import numpy as np

    krng = np.random.default_rng(1000)
    kx, ky = krng.random(1000), krng.random(1000)
    t_k = time.perf_counter()
    zk64 = prob.predict(theta, kx, ky)[0]
    t_k64 = time.perf_counter() - t_k
    mode = egs.PRECISION_MIXED_BF16 if args.plain_bf16 else egs.PRECISION_MIXED
    prob.set_precision(mode, args.band_fp64, args.band_fp32)
    prob.set_single_product(args.single_band, args.single_tau)
    prob.set_fp8_product(fp8_tau(args))
    if args.fp64_product_band is not None:
        prob.set_fp64_product_band(args.fp64_product_band)
    mixed = timed(egs, ctx, prob, theta, args.steps, args.warmup, "mixed")
    rel_err = abs(mixed["loglik"] - fp64["loglik"]) / abs(fp64["loglik"])
    t_k = time.perf_counter()
    zkmx = prob.predict(theta, kx, ky)[0]
    t_kmx = time.perf_counter() - t_k
    kriging = {"points": 1000, "max_abs_diff_vs_fp64": float(np.max(np.abs(zkmx - zk64))),
               "rel_l2_diff_vs_fp64": float(np.linalg.norm(zkmx - zk64) / np.linalg.norm(zk64)),
               "fp64_ms": 1000.0 * t_k64, "mixed_ms": 1000.0 * t_kmx}
    log("kriging of 1000 points, mixed vs FP64: max abs diff %.3g, rel L2 %.3g"
        % (kriging["max_abs_diff_vs_fp64"], kriging["rel_l2_diff_vs_fp64"]))

    # dominant kernel of the mixed run
    peaks = {"update_fp64": ("fp64", 1.0), "update_fp32": ("fp32", 1.0),
             "update_16": ("bf16", 1.0) if args.plain_bf16 else ("f16", 3.0), "update_16_select": ("f16", 1.0)}
    dom = max(peaks, key=lambda k: mixed["kernels"][k]["ms"])
    d = mixed["kernels"][dom]
    achieved = d["flops"] / (d["ms"] * 1e-3) / 1e12 if d["ms"] > 0 else 0.0
    prec, mfma_per_flop = peaks[dom]
    pc = mixed["product_counts"]
    nblk = pc["split"] + pc["single"] + pc.get("fp8", 0)
    if dom == "update_16_select" and nblk > 0:
        # FP16-MFMA time per product: 3 for split blocks, 1 for single-product
        # ones, 1/2 for FP8 ones (e4m3 MFMA at twice the FP16 rate)
        mfma_per_flop = (3.0 * pc["split"] + pc["single"] + 0.5 * pc.get("fp8", 0)) / nblk
    traffic, alg_bytes = measured_traffic(dom, args)
    dom64 = fp64["kernels"]["update_fp64"]
    ach64 = dom64["flops"] / (dom64["ms"] * 1e-3) / 1e12 if dom64["ms"] > 0 else 0.0

    out = {
        "metric": METRIC,
        "value": mixed["evals_per_s"],
        "unit": "llh evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32+bf16" if args.plain_bf16 else ("f64+f32+f16+fp8" if fp8_tau(args) > 0 else "f64+f32+f16"),
        "data": "synthetic: jittered grid locations (srand(0)), z = L(theta) e from the dlarnv stream, "
                "generated on the GPU in FP64",
        "config": {"workload": "univariate Matern likelihood, n=%d, theta=(1.0, 0.05, 0.5), dts=%d, mixed plan "
                               "FP64 band <=%d, FP32 band <=%d, %s beyond" % (
                                   args.n, args.dts, args.band_fp64, args.band_fp32,
                                   "BF16" if args.plain_bf16 else (
                                       "row-scaled split-FP16" if args.single_band < 0 else
                                       "row-scaled split-FP16, one FP16 product beyond band %d where the panel "
                                       "rows' scales multiply to <= %g of the panel's largest squared%s"
                                       % (args.single_band, args.single_tau,
                                          ", one FP8 (e4m3) product for FP32 tiles where they multiply to <= %g"
                                          % fp8_tau(args) if fp8_tau(args) > 0 else ""))),
                   "n": args.n, "dts": args.dts, "band_fp64": args.band_fp64, "band_fp32": args.band_fp32,
                   "single_band": args.single_band, "single_tau": tau_json(args),
                   "fp64_product_band": product_band(args, bivariate=False),
                   "fp8_tau": fp8_tau(args),
                   "operands": operands(args)},
        "roofline": {
            "bound": "mfma",
            "kernel": dom,
            "achieved": achieved,
            "peak": PEAK[prec],
            "unit": "TFLOP/s",
            "frac": achieved / PEAK[prec],
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes": alg_bytes,
            "traffic_over_algorithmic": (traffic / alg_bytes) if traffic and alg_bytes else None,
            "launches": d["launches"],
            "avg_launch_ms": d["ms"] / max(1, d["launches"]),
            "mfma_per_product": mfma_per_flop,
            "frac_mfma_issued": achieved * mfma_per_flop / PEAK[prec],
            "note": "achieved and frac count algorithmic flops (2 ts^3 per tile update) over the launch time; "
                    "frac_mfma_issued weights them by the FP16-MFMA time each product takes (3 per split product, "
                    "1 per single one, 1/2 per FP8 one, from the block counts of every per-block launch, FP32 tiles "
                    "and the FP64 band's together); peak is the FP16 dense peak; algorithmic_bytes is each "
                    "updated tile read and written once",
        },
        "fp64": {
            "evals_per_s": fp64["evals_per_s"], "ms_per_step": fp64["ms_per_step"], "steps": fp64_steps,
            "loglik": fp64["loglik"],
            "whole_eval_tflops": args.n ** 3 / 3.0 / (fp64["ms_per_step"] * 1e-3) / 1e12,
            "update_kernel": {"achieved": ach64, "peak": PEAK["fp64"], "frac": ach64 / PEAK["fp64"]},
            "phases_ms_per_step": fp64["phases_ms_per_step"],
            "kernels_per_step": kernel_table(fp64["kernels"], 1),
        },
        "mixed": {
            "loglik": mixed["loglik"], "rel_llh_err_vs_fp64": rel_err,
            "logdet_err_vs_fp64": mixed["logdet"] - fp64["logdet"],
            "quad_rel_err_vs_fp64": (mixed["quad"] - fp64["quad"]) / fp64["quad"],
            "refinement": mixed["refinement"],
            "kriging_at_theta": kriging,
            "block_updates_beyond_single_band": mixed["product_counts"],
            "speedup_vs_fp64": mixed["evals_per_s"] / fp64["evals_per_s"],
            "phases_ms_per_step": mixed["phases_ms_per_step"],
            "kernels_per_step": kernel_table(mixed["kernels"], 1),
        },
    }
    prob.close()
    ctx.close()
    if not args.no_serve:
        out["serve"] = serve_phase(egs, args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.n, args.cpu_sample, args.dts, theta, args.cpu_reps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
