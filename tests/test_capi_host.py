"""CPU-side checks of the C ABI: the library loads, exports every entry point
declared in include/*.h, and its host-only logic matches the oracle."""
import glob
import os
import re

import numpy as np
import pytest

import exageostatcpp_amd as egs
from exageostatcpp_amd import capi
from oracle import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(egs_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 20
    L = capi.lib()
    for n in sorted(names):
        assert hasattr(L, n), n
    assert names == set(capi.EXPORTS), names ^ set(capi.EXPORTS)


def test_no_device_is_a_clean_error():
    if os.path.exists("/dev/kfd") and os.environ.get("HIP_VISIBLE_DEVICES", "") != "-1":
        pytest.skip("a GPU may be present")
    with pytest.raises(egs.EGSError) as e:
        egs.Context(0)
    assert e.value.code in (capi.ERR_NO_DEVICE, capi.ERR_HIP)


@pytest.mark.parametrize("n,seed", [(9, 0), (16, 0), (1600, 0), (40000, 3), (1001, 17)])
def test_generate_locations_matches_oracle(n, seed):
    x, y = egs.generate_locations(n, seed=seed)
    xo, yo = orc.generate_locations(n, seed=seed)
    assert np.array_equal(x, xo) and np.array_equal(y, yo)


def test_generate_locations_3d_matches_oracle():
    x, y, z = egs.generate_locations(125, dim=3, seed=2)
    xo, yo, zo = orc.generate_locations(125, seed=2, dim=3)
    assert np.array_equal(x, xo) and np.array_equal(y, yo) and np.array_equal(z, zo)


def test_invalid_arguments_are_rejected():
    with pytest.raises(egs.EGSError):
        egs.generate_locations(0)
