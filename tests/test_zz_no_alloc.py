"""rewl_thermodynamics_noalloc, the allocation-free variant of
rewl_thermodynamics (static device scratch): called through the C ABI, its
results equal the allocating entry point's bit for bit on two calls in a row,
the second smaller than the first, and it refuses sizes beyond its scratch."""
import ctypes as C
import os

import numpy as np
import pytest

from rewl_simulator_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def exact_curve(L):
    f = np.load(os.path.join(GOLDEN, f"ising_L{L}_lng.npz"))
    return f["E"].astype(np.float64), f["lng"]

pytestmark = pytest.mark.gpu


def _call(fn, E, lng, N, nT, t_min=0.01, t_max=50.01):
    dp = C.POINTER(C.c_double)
    E = np.ascontiguousarray(E, dtype=np.float64)
    lng = np.ascontiguousarray(lng, dtype=np.float64)
    out = np.full((nT, 5), np.nan)
    rc = fn(E.ctypes.data_as(dp), lng.ctypes.data_as(dp), len(E), N, t_min, t_max, nT, out.ctypes.data_as(dp))
    return rc, out


def test_noalloc_thermodynamics_equals_allocating_entry_point(gpu):
    lib = abi.load_engine()
    cases = []
    # first call: the exact L = 64 Ising curve (2049 levels), 50 000 temperatures
    E64, lng64 = exact_curve(64)
    cases.append((E64, lng64, 64 * 64, 50000))
    # second call, smaller: the exact L = 16 curve, 1000 temperatures
    E16, lng16 = exact_curve(16)
    cases.append((E16, lng16, 16 * 16, 1000))
    for k, (E, lng, N, nT) in enumerate(cases):
        print(f"case: call {k}, {len(E)} bins, {nT} temperatures", flush=True)
        rc_a, ref = _call(lib.rewl_thermodynamics, E, lng, N, nT)
        rc_b, got = _call(lib.rewl_thermodynamics_noalloc, E, lng, N, nT)
        assert rc_a == abi.REWL_OK and rc_b == abi.REWL_OK, abi.load_engine().rewl_last_error()
        assert np.isfinite(ref).all()
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_noalloc_thermodynamics_refuses_sizes_beyond_its_scratch(gpu):
    lib = abi.load_engine()
    E = np.linspace(-1.0, 1.0, 16385)
    rc, _ = _call(lib.rewl_thermodynamics_noalloc, E, np.zeros_like(E), 100, 10)
    assert rc == abi.REWL_ERR_INVALID
    assert b"16384" in lib.rewl_last_error()
    rc, _ = _call(lib.rewl_thermodynamics_noalloc, E[:100], np.zeros(100), 100, 65537)
    assert rc == abi.REWL_ERR_INVALID
