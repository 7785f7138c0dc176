"""The reference's random-field generator restated in Python: TEST
INFRASTRUCTURE ONLY (imported by tests/ as the checker, never by the product).

generate_random_field (random_fields.hpp:41-75) draws, for every site, 20
discarded values and then the site's field from one std::mt19937_64 engine,
seeded from the clock in the reference (random_fields.hpp:48) and explicitly
here; a Gaussian value is get_gaussian's Box-Muller step with the rejection
u <= epsilon and a uniform coin choosing the cos or sin branch
(random_number_generators.hpp:33-57), a uniform value 2h(-1 + 2u)
(random_fields.hpp:28-33).  Uniform doubles are libstdc++'s
uniform_real_distribution<double>(0, 1) on a 64-bit engine: one engine call,
double(x) / 2^64, kept below 1.
"""
import math

import numpy as np

DISORDER_UNIFORM = 0
DISORDER_GAUSSIAN = 1

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
        if distribution == DISORDER_UNIFORM:
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
