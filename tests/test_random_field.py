"""The reference's random-field generator (generate_random_field,
random_fields.hpp:41-75; get_gaussian, random_number_generators.hpp:33-57)
restated in the engine library with an explicit seed, pinned against a
pure-Python restatement of the same steps on a pure-Python mt19937_64 (itself
pinned by the ISO C++ known answer) and by the distributions' moments.  Runs
on the host only."""
import math

import numpy as np
import pytest

from rewl_simulator_amd import abi, configs

MASK = (1 << 64) - 1


class MT19937_64:
    """std::mt19937_64 (ISO C++ [rand.predef]: w=64, n=312, m=156, r=31)."""

    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & MASK
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & MASK
        self.i = 312

    def __call__(self):
        if self.i >= 312:
            for k in range(312):
                y = (self.mt[k] & 0xFFFFFFFF80000000) | (self.mt[(k + 1) % 312] & 0x7FFFFFFF)
                v = self.mt[(k + 156) % 312] ^ (y >> 1)
                if y & 1:
                    v ^= 0xB5026F5AA96619E9
                self.mt[k] = v
            self.i = 0
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & MASK


def canonical(x):
    """uniform_real_distribution<double>(0, 1) on a 64-bit engine
    (libstdc++ generate_canonical: one draw, double(x) / 2^64, kept below 1)."""
    u = float(x) / 18446744073709551616.0
    return u if u < 1.0 else math.nextafter(1.0, 0.0)


def field_py(seed, distribution, strength, n):
    g = MT19937_64(seed)
    eps = np.finfo(np.float64).eps
    two_pi = 2.0 * math.acos(-1.0)

    def draw():
        if distribution == abi.DISORDER_UNIFORM:
            return 2. * strength * (-1. + 2. * canonical(g()))
        u, v = canonical(g()), canonical(g())
        while u <= eps:
            u, v = canonical(g()), canonical(g())
        mag = strength * math.sqrt(-2.0 * math.log(u))
        z0 = mag * math.cos(two_pi * v) + 0.
        z1 = mag * math.sin(two_pi * v) + 0.
        rc = canonical(g()) < 0.5
        return z0 * float(rc) + z1 * float(not rc)

    out = []
    for _ in range(n):
        for _ in range(20):
            draw()
        out.append(draw())
    return np.array(out)


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
