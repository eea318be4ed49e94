"""Writes mle_n1600.json and mle_c2_n1936.json: maximum-likelihood estimates
on the oracle.

mle_n1600.json: 40 x 40 jittered grid (srand(0)), z = L(theta) e with
theta = (1, 0.1, 0.5), dts 320, seed 0 -- the reference's synthetic generator
restated by the oracle.

mle_c2_n1936.json: the smooth-field workflow at a size the oracle finishes in
seconds: theta = (1, 0.1, 1.0), a 44 x 44 jittered grid (srand(0)) and 100
uniform held-out points (numpy seed 2036) whose observations are generated
jointly with the grid's (z = L(theta) e over all 2036 points, dts 320, seed 0);
the estimate uses the 1936 grid observations only.

Each optimum is found with scipy's Nelder-Mead on the oracle's FP64
log-likelihood (an optimiser independent of the library's BOBYQA), restarted
until the simplex no longer moves, and checked by a central-difference gradient.
Run from the repository root: python tests/golden/make_mle_fixture.py
"""
import json
import os
import sys

import numpy as np
from scipy.optimize import minimize

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import orc  # noqa: E402

N, DTS, THETA, LB, UB = 1600, 320, [1.0, 0.1, 0.5], [0.01] * 3, [5.0] * 3
C2_GRID, C2_HELD, C2_THETA = 1936, 100, [1.0, 0.1, 1.0]


def c2_data():
    """Grid and held-out locations of the smooth-field case with their jointly
    generated observations: ((x, y), z_obs, (xm, ym), z_held)."""
    x, y = orc.generate_locations(C2_GRID, seed=0)
    rng = np.random.default_rng(C2_GRID + C2_HELD)
    xm, ym = rng.random(C2_HELD), rng.random(C2_HELD)
    z = orc.generate_observations(0, (np.concatenate([x, xm]), np.concatenate([y, ym])), C2_THETA, DTS, seed=0)
    return (x, y), z[:C2_GRID], (xm, ym), z[C2_GRID:]


def fit(x, y, z, theta0):
    def f(v):
        if np.any(v < LB) or np.any(v > UB):
            return 1e30
        return -orc.loglik(0, (x, y), list(v), z, DTS)["llh"]

    v = np.array(theta0)
    for _ in range(4):
        r = minimize(f, v, method="Nelder-Mead", options=dict(xatol=1e-9, fatol=1e-11, maxiter=4000))
        moved = np.max(np.abs(r.x - v))
        v = r.x
        if moved < 1e-8:
            break
    h = 1e-5
    grad = [(f(v + h * e) - f(v - h * e)) / (2 * h) for e in np.eye(3)]
    return v, -f(v), grad


def write(name, out):
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), name), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def main():
    x, y = orc.generate_locations(N, seed=0)
    z = orc.generate_observations(0, (x, y), THETA, DTS, seed=0)
    v, llh, grad = fit(x, y, z, THETA)
    write("mle_n1600.json", {"_about": "maximum-likelihood estimate of the n = 1600 case (make_mle_fixture.py)",
                             "n": N, "dts": DTS, "theta_true": THETA, "lb": LB, "ub": UB,
                             "theta_hat": list(map(float, v)), "loglik": float(llh),
                             "gradient": list(map(float, grad))})
    (x, y), zo, (xm, ym), zh = c2_data()
    v, llh, grad = fit(x, y, zo, C2_THETA)
    zm, mspe = orc.krige(0, (x, y), (xm, ym), list(v), zo, DTS, z_actual=zh)
    write("mle_c2_n1936.json", {"_about": "smooth-field estimate and held-out kriging at n = 1936 "
                                          "(make_mle_fixture.py)",
                                "n": C2_GRID, "held_out": C2_HELD, "dts": DTS, "theta_true": C2_THETA, "lb": LB,
                                "ub": UB, "z_obs_sum": float(np.sum(zo)), "z_held_sum": float(np.sum(zh)),
                                "theta_hat": list(map(float, v)), "loglik": float(llh),
                                "gradient": list(map(float, grad)), "mspe_at_theta_hat": float(mspe)})


if __name__ == "__main__":
    main()
