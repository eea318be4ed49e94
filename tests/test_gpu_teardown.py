"""Process exit with library handles still open.

A program that ends without closing its contexts, problems or servers must
exit cleanly: the binding releases open handles at interpreter exit (servers,
then problems, then contexts) while the HIP runtime is still up, and a problem
whose context was finalized first frees nothing on the device when it is
destroyed.  Each case runs in a child process and must exit with status 0.
"""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PRELUDE = """
import sys
sys.path.insert(0, %r)
import exageostatcpp_amd as egs
x, y = egs.generate_locations(700, seed=3)
theta = [1.0, 0.1, 0.5]
""" % ROOT

CASES = {
    "context_and_problem_left_open": """
ctx = egs.Context(0)
p = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=192)
p.generate_observations(theta, seed=1)
print("llh", p.loglik(theta)[0], flush=True)
""",
    "mixed_problem_left_open": """
ctx = egs.Context(0)
p = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=192)
p.generate_observations(theta, seed=1)
p.set_precision(egs.PRECISION_MIXED)
print("llh", p.loglik(theta)[0], flush=True)
""",
    "context_closed_before_problem": """
ctx = egs.Context(0)
p = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=192)
p.generate_observations(theta, seed=1)
print("llh", p.loglik(theta)[0], flush=True)
ctx.close()
del ctx
p.close()
q = egs.Problem(egs.Context(0), "UnivariateMaternStationary", x, y, dts=192)
print("second problem created", flush=True)
""",
    "server_left_open": """
ctx = egs.Context(0)
p = egs.Problem(ctx, "UnivariateMaternStationary", x, y, dts=192)
z = p.generate_observations(theta, seed=1)
srv = egs.capi.Server(0, 1 << 30, 2)
r = srv.add_region(0, x, y, z, dts=192)
print("served", srv.wait(srv.submit_loglik(r, theta))[0], flush=True)
""",
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_exit_with_open_handles_is_clean(case):
    code = textwrap.dedent(PRELUDE) + textwrap.dedent(CASES[case])
    print("teardown case:", case, flush=True)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=90)
    assert r.returncode == 0, "case %s exited with %d\nstdout:\n%s\nstderr:\n%s" % (case, r.returncode, r.stdout,
                                                                                   r.stderr[-3000:])
    assert "llh" in r.stdout or "served" in r.stdout
