"""Writes mle_n1600.json: the maximum-likelihood estimate of the n = 1600 case.

Data: 40 x 40 jittered grid (srand(0)), z = L(theta) e with theta = (1, 0.1, 0.5),
dts 320, seed 0 -- the reference's synthetic generator restated by the oracle.
The optimum is found with scipy's Nelder-Mead on the oracle's FP64
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


def main():
    x, y = orc.generate_locations(N, seed=0)
    z = orc.generate_observations(0, (x, y), THETA, DTS, seed=0)

    def f(v):
        if np.any(v < LB) or np.any(v > UB):
            return 1e30
        return -orc.loglik(0, (x, y), list(v), z, DTS)["llh"]

    v = np.array(THETA)
    for _ in range(4):
        r = minimize(f, v, method="Nelder-Mead", options=dict(xatol=1e-9, fatol=1e-11, maxiter=4000))
        moved = np.max(np.abs(r.x - v))
        v = r.x
        if moved < 1e-8:
            break
    h = 1e-5
    grad = [(f(v + h * e) - f(v - h * e)) / (2 * h) for e in np.eye(3)]
    out = {"_about": __doc__.strip().splitlines()[0], "n": N, "dts": DTS, "theta_true": THETA, "lb": LB, "ub": UB,
           "theta_hat": list(map(float, v)), "loglik": float(-f(v)), "gradient": list(map(float, grad))}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "mle_n1600.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
