"""Pins of the CPU oracle: known-answer tests, closed forms, brute force.

None of these retype the oracle's own formulas: each compares it with a
standard's known answer, a fixture of the reference's own tests, an
independent enumeration, or a value the mathematics fixes.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from rewl_simulator_amd import abi, configs, exact

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mt19937_64_standard_kat():
    # ISO C++ [rand.predef]: the 10000th output of default-seeded mt19937_64
    assert int(O.mt19937_64(5489, 10000)[-1]) == 9981545732273789042


@pytest.mark.parametrize("ctr,key,expect", [
    # Random123 kat_vectors, philox4x32 R=10
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_random123_kat(ctr, key, expect):
    assert O.philox4x32_10(ctr, key) == expect


def test_neighbor_order_matches_reference_fixture():
    """The reference's tester fixes the table order (x-1,y),(x+1,y),(x,y-1),(x,y+1);
    the functor the Hamiltonians use orders them (x,y-1),(x-1,y),(x+1,y),(x,y+1)."""
    fx = json.load(open(os.path.join(HERE, "golden", "neighbors_3x3.json")))
    for site, (xm, xp, ym, yp) in enumerate(fx["table"]):
        assert O.neighbors(3, site) == [ym, xm, xp, yp]


def _libstdcxx_shuffle(words, n, base):
    """libstdc++'s std::shuffle for a 64-bit engine, restated: an even length
    first swaps position 1 with a draw in {0,1}; then positions are swapped in
    pairs from one uniform_int_distribution call over r*(r+1) values, each call
    reduced with Lemire's 128-bit multiply-and-reject."""
    it = iter(words)
    used = [0]

    def below(rng):
        x = next(it); used[0] += 1
        prod = x * rng
        lo = prod & ((1 << 64) - 1)
        if lo < rng:
            thr = ((1 << 64) - rng) % rng
            while lo < thr:
                x = next(it); used[0] += 1
                prod = x * rng
                lo = prod & ((1 << 64) - 1)
        return prod >> 64

    v = list(range(base, base + n))
    i = 1
    if n % 2 == 0:
        j = below(2)
        v[i], v[j] = v[j], v[i]
        i += 1
    while i != n:
        r = i + 1
        x = below(r * (r + 1))
        p1, p2 = x // (r + 1), x % (r + 1)
        v[i], v[p1] = v[p1], v[i]; i += 1
        v[i], v[p2] = v[p2], v[i]; i += 1
    return v, used[0]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 31, 32, 64, 255])
@pytest.mark.parametrize("seed", [1, 7, 20240501])
def test_shuffle_restatement_matches_libstdcxx(n, seed):
    perm, draws = O.shuffle_mt(seed, n, 100)
    words = [int(x) for x in O.mt19937_64(seed, 4 * n + 16)]
    mine, used = _libstdcxx_shuffle(words, n, 100)
    assert perm == mine
    assert draws == used


def _bond_energy(model, L, spins, J=1.0, K=0.5, field=None):
    """Direct bond sum, every bond once (right and down neighbours)."""
    N = L * L
    if model == abi.ISING2D:
        s = spins.reshape(L, L).astype(float)
        E = -J * (s * np.roll(s, -1, 1)).sum() - J * (s * np.roll(s, -1, 0)).sum()
        if field is not None:
            E -= (field.reshape(L, L) * s).sum()
        return E
    sg = spins[0::2].reshape(L, L).astype(float)
    ta = spins[1::2].reshape(L, L).astype(float)
    E = 0.0
    for ax in (0, 1):
        ss = sg * np.roll(sg, -1, ax)
        tt = ta * np.roll(ta, -1, ax)
        E += (-J * (ss + tt) - K * ss * tt).sum()
    if field is not None:
        E -= (field.reshape(L, L) * sg * ta).sum()
    return E


@pytest.mark.parametrize("L", [4, 6, 16])
def test_ising_energy_closed_forms(L):
    N = L * L
    s = configs.ising_system(L, -2.0 * N, 2.0 * N + 4, 4.0, 1, 1)
    assert O.energy_of(s, np.ones(N, np.int8)) == -2.0 * N
    checker = np.array([1 if (i % L + i // L) % 2 == 0 else -1 for i in range(N)], np.int8)
    assert O.energy_of(s, checker) == 2.0 * N
    rng = np.random.default_rng(L)
    for _ in range(5):
        sp = rng.choice([-1, 1], size=N).astype(np.int8)
        assert O.energy_of(s, sp) == _bond_energy(abi.ISING2D, L, sp)


@pytest.mark.parametrize("L", [4, 8])
def test_ashkin_teller_energy_closed_forms(L):
    N = L * L
    s = configs.at_system(L, -5.0 * N, 3.0 * N + 8, 8.0, 1, 1)
    assert O.energy_of(s, np.ones(2 * N, np.int8)) == -5.0 * N          # -2N(2J+K)
    checker = np.array([1 if (i % L + i // L) % 2 == 0 else -1 for i in range(N)], np.int8)
    both = np.empty(2 * N, np.int8); both[0::2] = checker; both[1::2] = checker
    assert O.energy_of(s, both) == 3.0 * N                              # every bond +1.5
    rng = np.random.default_rng(L)
    for _ in range(5):
        sp = rng.choice([-1, 1], size=2 * N).astype(np.int8)
        assert O.energy_of(s, sp) == _bond_energy(abi.ASHKIN_TELLER2D, L, sp)


def test_random_field_energies_match_bond_sum():
    L = 6
    N = L * L
    h = configs.gaussian_strain(L, 0.5, seed=3)
    rng = np.random.default_rng(1)
    si = configs.ising_system(L, -100, 100, 8.0, 1, 1, field=h)
    sa = configs.at_system(L, -300, 300, 8.0, 1, 1, field=h)
    for _ in range(5):
        sp = rng.choice([-1, 1], size=N).astype(np.int8)
        assert O.energy_of(si, sp) == pytest.approx(_bond_energy(abi.ISING2D, L, sp, field=h), abs=1e-12)
        sp2 = rng.choice([-1, 1], size=2 * N).astype(np.int8)
        assert O.energy_of(sa, sp2) == pytest.approx(_bond_energy(abi.ASHKIN_TELLER2D, L, sp2, field=h), abs=1e-12)


@pytest.mark.parametrize("sysdesc,expect", [
    # reference defaults for L=16: [-512, 0), bin 8, 4 windows at 0.75
    (configs.ising_system(16, -512.0, 0.0, 8.0, 4, 1),
     [(-512, -216, 37), (-440, -144, 37), (-368, -72, 37), (-296, 0, 37)]),
    # full range [-512, 516), bin 4
    (configs.ising_full_range(16, 4.0, 4, 2), [(-512, 76, 147), (-368, 224, 148), (-220, 372, 148), (-72, 516, 147)]),
])
def test_window_construction_closed_form(sysdesc, expect):
    o = O.Oracle([sysdesc], configs.params())
    got = [o.window_geometry(0, k) for k in range(len(expect))]
    assert [(int(a), int(b), n) for a, b, _, n in got] == expect


def test_config2_window_geometry():
    s, p = configs.config2()
    o = O.Oracle(s, p)
    w0 = o.window_geometry(0, 0)
    w31 = o.window_geometry(0, 31)
    assert (w0[0], w0[1], w0[3]) == (-8192.0, -6312.0, 235)
    assert w31[1] == 8200.0      # top level E = 8192 included
    # snapping each start onto the previous window's grid leaves 235 or 236 bins
    assert {o.window_geometry(0, k)[3] for k in range(32)} <= {235, 236}


def test_recalculated_energy_invariant_after_sweeps():
    """Incremental energies must equal a fresh recalculation (clean models are
    integer-exact)."""
    for s in (configs.ising_full_range(8, 4.0, 3, 2),
              configs.at_system(6, -180.0, 116.0, 8.0, 3, 2)):
        o = O.Oracle([s], configs.params(lnf_initial=0.01))
        o.adjust_to_range()
        o.sweeps(200)
        for g in range(o.num_walkers):
            assert o.state(g)["E"] == o.recalculated_energy(g)


def slack_window_rfat():
    """A random-strain AT system (L = 8, bins of width 1.002) whose top window
    starts at -0.34: the reference's +0.5 range test admits energies down to
    -1.5 there, one bin width below the window minimum."""
    from rewl_simulator_amd import warmstart
    h = configs.gaussian_strain(8, 0.5, 50002)
    e0, e1 = configs.rfat_range(8, h, K=0.5)
    bs = (e1 - e0) / 511
    s = configs.at_system(8, e0, e0 + bs * 512, bs, 8, 4, K=0.5, field=h, seed=50002)
    return [s], warmstart.rfat_params(1e-4)


def test_every_attempt_counts_in_a_window_bin():
    """Each attempt increments exactly one bin of the walker's window, also for
    a proposal in the range test's slack below a window starting near 0 (whose
    quotient (E - min)/bs is below -1): without exchanges, the counts of a
    walker sum to sweeps x degrees of freedom."""
    systems, params = slack_window_rfat()
    o = O.Oracle(systems, params)
    geo = o.window_geometry(0, 7)
    assert -0.5 < geo[0] < 0 and geo[2] < 1.5
    o.adjust_to_range()
    sweeps = 300
    o.sweeps(sweeps)
    ndof = 2 * 8 * 8
    for g in range(o.num_walkers):
        c = o.counts(g).astype(np.int64) + o.visits(g).astype(np.int64)
        assert c.sum() == sweeps * ndof, g
        assert o.state(g)["E"] == o.recalculated_energy(g) or abs(o.state(g)["E"] - o.recalculated_energy(g)) < 1e-9
    o.close()


def test_oracle_wang_landau_converges_to_exact_L4():
    """Single window over the whole L=4 spectrum, the reference's own defaults
    (1/t schedule, all-visited flatness): ln g converges to the
    brute-force-pinned exact density of states."""
    E, g = exact.ising_dos(4)
    ref = {int(e): v for e, v in zip(E, g)}
    s = configs.ising_full_range(4, 4.0, 1, 1)
    for seed in (1, 2, 3):
        s["seed"] = seed
        o = O.Oracle([s], configs.params(lnf_final=1e-6, exchange=0))
        o.adjust_to_range()
        assert o.run(100000)
        lng = o.logdos(0)
        bins = [-32 + 4 * b for b in range(lng.size)]
        keep = [b for b, e in enumerate(bins) if ref.get(e, 0) > 0]
        est = lng[keep] - lng[keep[0]] + np.log(2.0)
        tru = np.array([np.log(ref[bins[b]]) for b in keep])
        assert np.max(np.abs(est - tru)) < 0.05


def test_stitch_and_normalise_exact_input():
    """Stitching windows cut from the exact curve and normalising recovers it."""
    s = configs.ising_full_range(8, 4.0, 3, 1)
    o = O.Oracle([s], configs.params())
    E, g = exact.ising_dos(8)
    lng_of = {int(e): exact._ln_int(v) for e, v in zip(E, g) if v > 0}
    for k in range(3):
        emin, emax, bs, nb = o.window_geometry(0, k)
        lw = np.array([lng_of.get(int(emin + bs * b), 0.0) + 3.0 * k for b in range(nb)])
        o.set_logdos(k, lw)
    Es, ls = o.stitched(0, states_per_site=2.0)
    for e, v in zip(Es, ls):
        if int(e) in lng_of and int(e) not in (-128 + 4, 128 - 4):
            assert v == pytest.approx(lng_of[int(e)], abs=1e-9)


def test_thermodynamics_two_level_closed_form():
    """Two bins E = {0, 1}, ln g = {0, 0}: both are end bins of the trapezoid,
    so Z' = (1 + e^{-1/T})/2 and U = e^{-1/T} / (1 + e^{-1/T}) (N = 1)."""
    T = np.array([0.5, 1.0, 3.0])
    out = O.thermodynamics(np.array([0.0, 1.0]), np.array([0.0, 0.0]), 1, tmin=0.5, tmax=3.5, nT=3)
    for t, row in zip([0.5, 1.5, 2.5], out):
        x = np.exp(-1.0 / t)
        assert row[1] == pytest.approx(x / (1 + x), rel=1e-14)
        assert row[2] == pytest.approx(x / (1 + x), rel=1e-14)
        assert row[0] == pytest.approx(-t * np.log((1 + x) / 2), rel=1e-13)
        assert row[4] == pytest.approx((x / (1 + x) - (x / (1 + x)) ** 2) / t ** 2, rel=1e-12)


def test_config4_ranges_and_visited_flatness():
    """Random-strain AT ranges: E(all-up) and E(checkerboard) from the bond sum;
    the oracle runs the visited-support flatness mode to flat events."""
    systems, params = configs.config4(realisations=2, sizes=(4, 6), bins=32, windows=2, replicas=2, lnf_final=1e-2)
    for s in systems:
        L = s["L"]
        up = np.ones(2 * L * L, np.int8)
        assert O.energy_of(s, up) == pytest.approx(s["e_min"], abs=1e-9)
    o = O.Oracle(systems, dict(params, sweeps_per_check=50, sweeps_per_exchange=25, flatness=0.5))
    o.adjust_to_range()
    assert o.run(2000)
    assert o.stats()["flat_events"] > 0


@pytest.mark.parametrize("average", [0, 1])
def test_window_average_weights(average):
    """average_and_redistribute_window: after a flat check every replica of a
    window holds sum_r w_r ln g_r (then REDUCE_LOGDOS), with w_r = c_r / sum c
    (the reference, rewl_simulation.hpp:393-397) or 1/R (its alternative,
    line 401).  Computed here from the pre-check ln g and counts in numpy."""
    s = configs.ising_full_range(6, 4.0, 2, 3)
    p = configs.params(flatness_mode=abi.FLAT_ALL_VISITED, schedule=abi.SCHED_HALVE, lnf_initial=0.5,
                       average_mode=average, reduce_logdos=0, exchange=0)
    o = O.Oracle([s], p)
    o.adjust_to_range()
    o.sweeps(400)
    R = 3
    for k in range(2):
        lng = np.array([o.logdos(k * R + r) for r in range(R)])
        cnt = np.array([o.counts(k * R + r) for r in range(R)], dtype=np.float64)
        if average == 0:
            tot = cnt.sum(0)
            want = np.zeros(lng.shape[1])
            for r in range(R):
                want = want + np.where(tot > 0, cnt[r] / np.where(tot > 0, tot, 1) * lng[r], 0)
            keep = tot == 0
        else:
            want = sum(lng[r] / R for r in range(R))
            keep = np.zeros(lng.shape[1], bool)
        lng_k = lng, want, keep
        if k == 0:
            first = lng_k
        else:
            second = lng_k
    o.check()
    for k, (lng, want, keep) in enumerate((first, second)):
        assert o.window_state(0, k)["iteration"] == 2, "the window was not flat"
        for r in range(R):
            got = o.logdos(k * R + r)
            np.testing.assert_allclose(got[~keep], want[~keep], rtol=1e-14, atol=1e-14)
            np.testing.assert_array_equal(got[keep], lng[r][keep])
