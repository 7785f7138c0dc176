"""The reference's random-field generator (generate_random_field,
random_fields.hpp:41-75; get_gaussian, random_number_generators.hpp:33-57)
restated in the engine library with an explicit seed, pinned against a
pure-Python restatement of the same steps on a pure-Python mt19937_64 (itself
pinned by the ISO C++ known answer) and by the distributions' moments.  Runs
on the host only."""
import math

import numpy as np
import pytest

from oracle.random_field import MT19937_64, field_py
from rewl_simulator_amd import abi, configs


def test_python_mt19937_64_iso_known_answer():
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


@pytest.mark.parametrize("distribution", [abi.DISORDER_UNIFORM, abi.DISORDER_GAUSSIAN])
@pytest.mark.parametrize("seed", [1, 42, 1700000000123456789])
def test_library_field_equals_restatement(seed, distribution):
    lib = configs.reference_random_field(4, 0.75, seed, distribution)
    py = field_py(seed, distribution, 0.75, 16)
    np.testing.assert_array_equal(lib, py)


def test_field_moments():
    g = configs.reference_random_field(128, 0.5, 7, abi.DISORDER_GAUSSIAN)
    n = g.size
    assert abs(g.mean()) < 4 * 0.5 / math.sqrt(n)
    assert g.std() == pytest.approx(0.5, rel=0.03)
    assert abs(((g / 0.5) ** 4).mean() - 3.0) < 0.15   # Gaussian kurtosis
    u = configs.reference_random_field(128, 0.5, 7, abi.DISORDER_UNIFORM)
    assert u.min() >= -1.0 and u.max() <= 1.0
    assert u.std() == pytest.approx(2 * 0.5 / math.sqrt(3.0), rel=0.03)   # U(-2s, 2s)


def test_rejects_unknown_distribution():
    import ctypes as C
    lib = abi.load_engine()
    out = np.empty(4)
    assert lib.rewl_random_field(C.c_uint64(1), 7, 1.0, 4, out.ctypes.data_as(C.POINTER(C.c_double))) == abi.REWL_ERR_INVALID
