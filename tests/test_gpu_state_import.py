"""Importing configurations into a context: the per-system import equals one
import per walker, and a failed range adjustment does not poison the next one.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rewl_simulator_amd import abi, configs
from rewl_simulator_amd.engine import Engine

pytestmark = pytest.mark.gpu


def test_system_import_equals_per_walker_import(gpu):
    """rewl_set_system_spins (import_DoFs after the broadcast, main.cpp:194-210)
    leaves every walker of the system with the given spins and their energy,
    and the other systems untouched; after the adjustment the chains equal
    the oracle's, which imports walker by walker."""
    print("case: system spins import", flush=True)
    s0 = configs.at_system(8, -5.0 * 64, 3.0 * 64 + 8.0, 8.0, 3, 2)
    s1 = configs.ising_full_range(6, 4.0, 2, 2, seed=5)
    p = configs.params(lnf_initial=0.05, adjust_mode=1)
    eng, ora = Engine([s0, s1], p), Oracle([s0, s1], p)
    rng = np.random.default_rng(11)
    spins = rng.choice(np.array([-1, 1], dtype=np.int8), size=2 * 64)
    eng.set_system_spins(0, spins)
    E = configs.at_energy(8, spins[0::2], spins[1::2])
    for g in range(6):
        np.testing.assert_array_equal(eng.spins(g), spins)
        assert eng.energy(g) == E
        ora.set_spins(g, spins)
    for g in range(6, eng.num_walkers):
        assert eng.energy(g) == -2.0 * 36   # all up, untouched
    with pytest.raises(abi.REWLError) as e:
        eng.sweeps(1)   # the imported states may lie outside their windows
    assert e.value.code == abi.REWL_ERR_RANGE
    eng.adjust_to_range(); ora.adjust_to_range()
    eng.sweeps(20); ora.sweeps(20)
    for g in range(eng.num_walkers):
        assert eng.energy(g) == ora.state(g)["E"], g
        np.testing.assert_array_equal(eng.spins(g), ora.spins(g))
    with pytest.raises(abi.REWLError):
        eng.set_system_spins(2, spins)
    with pytest.raises(abi.REWLError):
        eng.set_system_spins(1, spins)   # wrong length for L = 6 Ising
    eng.close()


def test_failed_adjust_then_reachable_state_adjusts(gpu):
    """The reference's range scan flips sigma only on Ashkin-Teller (its flavour
    toggles twice per wrap), so from all-up it reaches E <= N at most: a
    system whose range lies above N cannot be brought into its windows and
    the adjustment fails.  Once that system's walkers start from tau in a
    checkerboard (E = N, and sigma flips lead up to 3N) the next adjustment
    succeeds: the range error of the first call does not carry over."""
    print("case: failed adjust, import, adjust again", flush=True)
    L, N = 10, 100
    low = configs.at_system(L, -5.0 * N, -1.0 * N, 8.0, 3, 2)
    high = configs.at_system(L, 1.2 * N, 2.2 * N, 8.0, 2, 2, seed=9)
    eng = Engine([low, high], configs.params(adjust_mode=0, lnf_initial=0.05))
    with pytest.raises(abi.REWLError) as e:
        eng.adjust_to_range()
    assert e.value.code == abi.REWL_ERR_RANGE
    with pytest.raises(abi.REWLError):
        eng.sweeps(1)
    chk = np.ones(2 * N, dtype=np.int8)
    chk[1::2] = [1 if (i % L + i // L) % 2 == 0 else -1 for i in range(N)]
    assert configs.at_energy(L, chk[0::2], chk[1::2]) == float(N)
    eng.set_system_spins(1, chk)
    eng.adjust_to_range()
    eng.sweeps(10)
    for g in range(eng.num_walkers):
        w = eng.windows[eng.walker_window[g]]
        assert w["e_min"] - 0.5 <= eng.energy(g) < w["e_max"] + 0.5, g
    eng.close()


def test_set_logdos_round_trip_leaves_other_walkers(gpu):
    """rewl_set_logdos writes exactly one walker's column of the [bin][lane]
    histogram layout (through a staging copy and a scatter launch): it reads
    back bit for bit, every other walker keeps its ln g, and a second call
    with a longer window reuses or grows the staging buffer."""
    print("case: set_logdos round trip", flush=True)
    systems, params = configs.config1()
    with Engine(systems + [configs.ising_full_range(24, 8.0, 6, 3, seed=9)], params) as eng:
        eng.adjust_to_range()
        eng.sweeps(50)
        before = [eng.logdos(g) for g in range(eng.num_walkers)]
        rng = np.random.default_rng(5)
        # a walker of the second system (129 bins per window), then the first
        # and last walker of the first (147 bins: the staging buffer grows)
        picks = [eng.num_walkers - 2, 0, systems[0]["num_windows"] * systems[0]["replicas_per_window"] - 1]
        new = {}
        for g in picks:
            new[g] = rng.normal(0.0, 50.0, size=before[g].size)
            eng.set_logdos(g, new[g])
        for g in range(eng.num_walkers):
            got = eng.logdos(g)
            np.testing.assert_array_equal(got, new[g] if g in new else before[g])
        eng.sweeps(5)   # the engine still runs on the imported ln g
