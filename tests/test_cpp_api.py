"""The C++ host API (exageostatcpp_amd/api) driven by tests/cpp/test_exageostat_api.cpp.

The program's cases follow the reference's API tests (TestExaGeoStatApi.cpp,
TestPrediction.cpp, TestAllRFunctions.cpp); the cpu cases need no GPU.
"""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin", "test_exageostat_api")


def _run(*args, timeout=300):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/_bin/test_exageostat_api is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN] + list(args), capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    return r.stdout


def test_cpp_api_host_cases():
    out = _run("cpu")
    assert "cpu: 0 failure(s)" in out


@pytest.mark.gpu
def test_cpp_api_reference_fixtures_on_gpu():
    out = _run("gpu")
    assert "gpu: 0 failure(s)" in out


@pytest.mark.gpu
def test_cpp_api_mle_n1600_matches_oracle_optimum():
    with open(os.path.join(ROOT, "tests", "golden", "mle_n1600.json")) as fh:
        f = json.load(fh)
    th = f["theta_hat"]
    out = _run("mle", "%.12f" % th[0], "%.12f" % th[1], "%.12f" % th[2], "%.12f" % f["loglik"], timeout=600)
    assert "mle: 0 failure(s)" in out
