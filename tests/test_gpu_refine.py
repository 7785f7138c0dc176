"""rewl_refine on the GPU: the same chains and the same corrected ln g as the
oracle's restatement, and ln g within 0.01 of the exact density of states
after the refinement (the accuracy Wang-Landau alone does not reach)."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from rewl_simulator_amd import abi, configs
from rewl_simulator_amd.engine import Engine

from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
def test_refine_matches_oracle(gpu, rng):
    print(f"case: refine parity rng={rng}", flush=True)
    s = configs.ising_full_range(8, 4.0, 3, 3)
    p = configs.params(rng_mode=rng, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE,
                       lnf_final=1e-2, sweeps_per_check=100, sweeps_per_exchange=50)
    eng, ora = Engine([s], p), Oracle([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    eng.run(200); ora.run(200)
    eng.refine(3000, 50); ora.refine(3000, 50)
    for g in range(eng.num_walkers):
        assert eng.energy(g) == ora.state(g)["E"], g
        np.testing.assert_array_equal(eng.spins(g), ora.spins(g))
        np.testing.assert_array_equal(eng.visits(g), ora.visits(g))
        # device log against glibc log: equal to the last bit or one ulp
        np.testing.assert_allclose(eng.logdos(g), ora.logdos(g), rtol=4e-16, atol=4e-15)
    for k in range(eng.num_windows):
        assert eng.window_info(k)["lnf"] == 0.0
    eng.close(); ora.close()


def test_refine_in_parts_equals_one_call(gpu):
    """rewl_refine_begin + rewl_refine_sweeps in parts that cut across the
    exchange cadence + rewl_refine_end: the same chains and ln g as one
    rewl_refine call, bit for bit."""
    print("case: refine in parts", flush=True)
    s = configs.ising_full_range(8, 4.0, 3, 3)
    p = configs.params(flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE,
                       lnf_final=1e-2, sweeps_per_check=100, sweeps_per_exchange=50)
    one, parts = Engine([s], p), Engine([s], p)
    for e in (one, parts):
        e.adjust_to_range()
        e.run(200)
    one.refine(3000, 50)
    seen = []
    parts.refine(3000, 50, chunk=700, progress=seen.append)
    assert seen == [700, 1400, 2100, 2800, 3000]
    for g in range(one.num_walkers):
        assert one.energy(g) == parts.energy(g)
        np.testing.assert_array_equal(one.spins(g), parts.spins(g))
        np.testing.assert_array_equal(one.visits(g), parts.visits(g))
        np.testing.assert_array_equal(one.logdos(g).view(np.uint64), parts.logdos(g).view(np.uint64))
    one.close(); parts.close()


def test_refined_lng_within_0_01_of_exact_L16(gpu):
    """2D Ising L = 16 (the first workload's lattice: full spectrum, bins of
    4, 4 windows at 75% overlap), here 128 walkers per window: Wang-Landau to
    ln f < 1e-4, then 10^6 sweeps of refinement; the stitched curve (overlap
    join), normalised to 2^N, within 0.01 of the exact ln g at every level.

    The refined error is statistical: with 8 walkers per window and 4 x 10^5
    sweeps it is 0.04 (max over the 257 levels); 40 x the samples divide it
    by about 6."""
    print("case: refined accuracy L=16", flush=True)
    s = configs.ising_full_range(16, 4.0, 4, 128)
    p = configs.params(flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE, lnf_final=1e-4,
                       sweeps_per_check=1000, sweeps_per_exchange=1000, average_mode=1)
    f = np.load(os.path.join(GOLDEN, "ising_L16_lng.npz"))
    exact = {int(e): v for e, v in zip(f["E"], f["lng"])}
    with Engine([s], p) as eng:
        eng.adjust_to_range()
        eng.run(100000)
        eng.finalize()
        E0, l0 = eng.stitched(0, 2.0, join="overlap")
        eng.refine(1000000, 1000)
        E1, l1 = eng.stitched(0, 2.0, join="overlap")

    def err(E, lng):
        d = [lng[i] - exact[int(E[i])] for i in range(len(E)) if int(E[i]) in exact and np.isfinite(lng[i])]
        assert len(d) == len(exact)
        return max(abs(x) for x in d)

    e0, e1 = err(E0, l0), err(E1, l1)
    print(f"max |ln g - exact|: Wang-Landau to 1e-4 {e0:.4f}, refined {e1:.4f}", flush=True)
    assert e1 <= 0.01
    assert e1 < e0
