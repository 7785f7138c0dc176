"""Bit-exact parity of the HIP engine with the CPU oracle.

Both sides consume the same random stream (the reference's per-walker
mt19937_64 in replay mode, or the engine's Philox positions) from the same
initial state, so spins, energies, ln g, counts, visits and stream cursors
must agree bit for bit after every step.
"""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from rewl_simulator_amd import abi, configs
from rewl_simulator_amd.engine import Engine

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_same(eng, ora, what="", cursors=True):
    assert eng.num_walkers == ora.num_walkers
    for g in range(eng.num_walkers):
        st = ora.state(g)
        assert eng.energy(g) == st["E"], f"{what}: energy of walker {g}"
        np.testing.assert_array_equal(eng.spins(g), ora.spins(g), err_msg=f"{what}: spins of walker {g}")
        np.testing.assert_array_equal(bits(eng.logdos(g)), bits(ora.logdos(g)), err_msg=f"{what}: ln g of walker {g}")
        np.testing.assert_array_equal(eng.counts(g).astype(np.uint64), ora.counts(g), err_msg=f"{what}: counts {g}")
        np.testing.assert_array_equal(eng.visits(g), ora.visits(g), err_msg=f"{what}: visits {g}")
        assert eng.flavor(g) == st["flavor"], f"{what}: flavour of walker {g}"
        if cursors:
            assert eng.rng_cursor(g) == st["draws"], f"{what}: stream cursor of walker {g}"


def window_states_equal(eng, ora):
    for k, w in enumerate(eng.windows):
        e = eng.window_info(k)
        o = ora.window_state(w["system"], w["window"])
        assert e["lnf"] == o["lnf"], f"ln f of window {k}"
        assert e["iteration"] == o["iteration"], f"iteration of window {k}"
        assert e["done"] == o["done"]
        assert e["one_over_t"] == o["one_over_t"]


def make_pair(systems, params):
    eng = Engine(systems, params)
    ora = Oracle(systems, params)
    return eng, ora


def small_ising(L=16, windows=4, replicas=2, **kw):
    return configs.ising_full_range(L, 4.0, windows, replicas, **kw)


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
def test_ising_sweeps_bit_exact(gpu, rng):
    print(f"case: ising L=16 sweeps rng={rng}", flush=True)
    s = small_ising()
    p = configs.params(rng_mode=rng, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8,
                       schedule=abi.SCHED_HALVE)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    assert_same(eng, ora, "after adjust", cursors=(rng == abi.RNG_REPLAY))
    for chunk in (1, 7, 50, 200):
        eng.sweeps(chunk); ora.sweeps(chunk)
        assert_same(eng, ora, f"after {chunk} sweeps", cursors=(rng == abi.RNG_REPLAY))


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
def test_ising_small_lnf_bit_exact(gpu, rng):
    """Many sweeps at a small ln f so that acceptances and draws are frequent."""
    print(f"case: ising L=8 small lnf rng={rng}", flush=True)
    s = configs.ising_full_range(8, 4.0, 2, 3)
    p = configs.params(rng_mode=rng, lnf_initial=1e-3)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    eng.sweeps(3000); ora.sweeps(3000)
    assert_same(eng, ora, "after 3000 sweeps", cursors=(rng == abi.RNG_REPLAY))


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
def test_ashkin_teller_sweeps_bit_exact(gpu, rng):
    print(f"case: AT L=8 sweeps rng={rng}", flush=True)
    N = 64
    s = configs.at_system(8, -5.0 * N, 3.0 * N + 8.0, 8.0, 3, 2)
    p = configs.params(rng_mode=rng, lnf_initial=0.05)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    assert_same(eng, ora, "after adjust", cursors=(rng == abi.RNG_REPLAY))
    for chunk in (1, 13, 300):
        eng.sweeps(chunk); ora.sweeps(chunk)
        assert_same(eng, ora, f"after {chunk} sweeps", cursors=(rng == abi.RNG_REPLAY))


@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
def test_random_field_fp64_path_bit_exact(gpu, model, rng):
    """RFIM / random-strain AT: FP64 energies in the reference's operation order."""
    print(f"case: random field model={model} rng={rng}", flush=True)
    L = 8
    N = L * L
    h = configs.gaussian_strain(L, 0.5, seed=7)
    if model == abi.ISING2D:
        s = configs.ising_system(L, -2.0 * N - 40, 2.0 * N + 40, 8.0, 3, 2, field=h)
    else:
        s = configs.at_system(L, -5.0 * N - 40, 3.0 * N + 40, 12.0, 3, 2, field=h)
    p = configs.params(rng_mode=rng, lnf_initial=0.1)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    assert_same(eng, ora, "after adjust", cursors=(rng == abi.RNG_REPLAY))
    for chunk in (1, 40, 200):
        eng.sweeps(chunk); ora.sweeps(chunk)
        assert_same(eng, ora, f"after {chunk} sweeps", cursors=(rng == abi.RNG_REPLAY))


def test_random_field_slack_below_window_bit_exact(gpu):
    """Random-strain AT with a window starting at -0.34 and bins of width
    1.002: proposals in the +0.5 range test's slack below it (quotient below
    -1) count in bin 0 on both sides."""
    from rewl_simulator_amd import warmstart
    print("case: random-strain AT L=8, slack below window 7", flush=True)
    h = configs.gaussian_strain(8, 0.5, 50002)
    e0, e1 = configs.rfat_range(8, h, K=0.5)
    bs = (e1 - e0) / 511
    systems = [configs.at_system(8, e0, e0 + bs * 512, bs, 8, 4, K=0.5, field=h, seed=50002)]
    assert -0.5 < Oracle(systems, warmstart.rfat_params(1e-4)).window_geometry(0, 7)[0] < 0
    p = warmstart.rfat_params(1e-4)
    eng, ora = make_pair(systems, p)
    eng.adjust_to_range(); ora.adjust_to_range()
    for chunk in (50, 250):
        eng.sweeps(chunk); ora.sweeps(chunk)
        assert_same(eng, ora, f"after {chunk} sweeps", cursors=False)
    for _ in range(10):
        eng.sweeps(250); eng.exchange(); eng.check()
        ora.sweeps(250); ora.exchange(); ora.check()
    assert_same(eng, ora, "after 10 check cycles", cursors=False)
    window_states_equal(eng, ora)


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
@pytest.mark.parametrize("replicas", [1, 2, 5])
def test_rewl_cycles_bit_exact(gpu, rng, replicas):
    """Sweeps, replica exchange every sweep and flatness checks, reference
    defaults (1/t schedule, all-visited flatness, REDUCE_LOGDOS)."""
    print(f"case: rewl cycles rng={rng} replicas={replicas}", flush=True)
    s = configs.ising_full_range(6, 4.0, 4, replicas)
    p = configs.params(rng_mode=rng, sweeps_per_check=25, sweeps_per_exchange=1)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    for cycle in range(40):
        for _ in range(25):
            eng.sweeps(1); ora.sweeps(1)
            eng.exchange(); ora.exchange()
        eng.check(); ora.check()
        if cycle % 8 == 0 or cycle == 39:
            assert_same(eng, ora, f"cycle {cycle}")
            window_states_equal(eng, ora)
    assert eng.stats()["exchanges_accepted"] == ora.stats()["exchanges_accepted"]


@pytest.mark.parametrize("rng,average", [(abi.RNG_REPLAY, 0), (abi.RNG_PHILOX, 0), (abi.RNG_PHILOX, 1)])
def test_halving_run_and_finalize_bit_exact(gpu, rng, average):
    """Whole run to completion with min >= 0.8 mean flatness and halving;
    count-weighted (reference) and equal-weight replica averaging."""
    print(f"case: full halving run rng={rng} average={average}", flush=True)
    s = configs.ising_full_range(4, 4.0, 2, 3)
    p = configs.params(rng_mode=rng, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8,
                       schedule=abi.SCHED_HALVE, lnf_final=1e-3, sweeps_per_check=100, sweeps_per_exchange=10,
                       average_mode=average)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    done_e = eng.run(400)
    done_o = ora.run(400)
    assert done_e == done_o
    assert_same(eng, ora, "end of run")
    window_states_equal(eng, ora)
    eng.finalize(); ora.finalize()
    assert_same(eng, ora, "after finalize")
    Ee, le = eng.stitched(0, 2.0)
    Eo, lo = ora.stitched(0, 2.0)
    np.testing.assert_array_equal(Ee, Eo)
    np.testing.assert_allclose(le, lo, rtol=0, atol=1e-12)


def test_neglog_band_covers_fp32_error(gpu):
    """The throughput kernel screens u < exp(-d) with T = -ln u in FP32 and an
    absolute band of 2^-12; every 32-bit Philox word must keep T well inside it."""
    import ctypes as C
    lib = abi.load_engine()
    err = C.c_double()
    abi.check(lib.rewl_selfcheck(0, C.byref(err)))
    print(f"max |T - (-ln u)| = {err.value:.3e}", flush=True)
    assert err.value < 2.0 ** -12 / 16


@pytest.mark.parametrize("cfg", ["config2", "config3"])
def test_production_configs_few_sweeps_bit_exact(gpu, cfg):
    """Full-size workloads (L=64 Ising, 2048 walkers; L=32 Ashkin-Teller, 512
    walkers): a few Philox sweeps and an exchange, bit-exact with the oracle."""
    print(f"case: {cfg} few sweeps", flush=True)
    systems, params = getattr(configs, cfg)()
    eng, ora = make_pair(systems, params)
    eng.adjust_to_range(); ora.adjust_to_range()
    eng.sweeps(3); ora.sweeps(3)
    eng.exchange(); ora.exchange()
    eng.sweeps(2); ora.sweeps(2)
    eng.check(); ora.check()
    rng = np.random.default_rng(0)
    for g in sorted(set(rng.integers(0, eng.num_walkers, 24).tolist()) | {0, eng.num_walkers - 1}):
        st = ora.state(g)
        assert eng.energy(g) == st["E"], g
        np.testing.assert_array_equal(eng.spins(g), ora.spins(g))
        np.testing.assert_array_equal(bits(eng.logdos(g)), bits(ora.logdos(g)))
        np.testing.assert_array_equal(eng.counts(g).astype(np.uint64), ora.counts(g))


def test_config4_full_size_bit_exact(gpu):
    """The batched random-strain Ashkin-Teller workload at its stated size:
    256 realisations (64 each of L = 16, 24, 32, 48), 8 windows x 4 walkers,
    512 bins, 8192 walkers in one context, so the field-model kernel runs in
    its production launch shape (one 32-walker realisation per workgroup,
    512-bin windows, the L = 48 field and bound ring in LDS, mixed L in one
    grid).  Sweeps, an exchange of each phase and two checks; the first and
    last walker of every realisation plus a random sample are compared bit
    for bit with the oracle, which runs each realisation with its own
    streams."""
    print("case: config4 full size (256 realisations, 8192 walkers)", flush=True)
    systems, params = configs.config4()
    assert len(systems) == 256 and sorted({s["L"] for s in systems}) == [16, 24, 32, 48]
    eng, ora = make_pair(systems, params)
    assert eng.num_walkers == 8192
    eng.adjust_to_range(); ora.adjust_to_range()
    eng.sweeps(3); ora.sweeps(3)
    eng.exchange(); ora.exchange()
    eng.sweeps(2); ora.sweeps(2)
    eng.check(); ora.check()
    eng.sweeps(4); ora.sweeps(4)
    eng.exchange(); ora.exchange()
    eng.sweeps(1); ora.sweeps(1)
    eng.check(); ora.check()
    per = 8 * 4
    sample = set()
    for r in range(len(systems)):
        sample |= {r * per, r * per + per - 1}
    sample |= set(np.random.default_rng(4).integers(0, eng.num_walkers, 64).tolist())
    for g in sorted(sample):
        st = ora.state(g)
        assert eng.energy(g) == st["E"], f"energy of walker {g} (L={systems[g // per]['L']})"
        assert eng.flavor(g) == st["flavor"], g
        np.testing.assert_array_equal(eng.spins(g), ora.spins(g), err_msg=f"spins of walker {g}")
        np.testing.assert_array_equal(bits(eng.logdos(g)), bits(ora.logdos(g)), err_msg=f"ln g of walker {g}")
        np.testing.assert_array_equal(eng.counts(g).astype(np.uint64), ora.counts(g), err_msg=f"counts {g}")
        np.testing.assert_array_equal(eng.visits(g), ora.visits(g), err_msg=f"visits {g}")
    window_states_equal(eng, ora)
    assert eng.stats()["exchanges_accepted"] == ora.stats()["exchanges_accepted"]
    eng.close()


@pytest.mark.parametrize("producers", [-1, 1, 2, 3])
@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
def test_fast_path_producers_bit_exact(gpu, producers, model):
    """The throughput kernel with the walkers drawing their own Philox words
    (producers = -1) or with 1-3 producer wavefronts filling the threshold
    ring: identical chains, bit-exact with the oracle."""
    print(f"case: fast path producers={producers} model={model}", flush=True)
    if model == abi.ISING2D:
        s = configs.ising_full_range(12, 8.0, 3, 7)
    else:
        s = configs.at_system(12, -5.0 * 144, 3.0 * 144 + 8.0, 8.0, 3, 5)
    p = configs.params(producers=producers, lnf_initial=0.02, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.5,
                       schedule=abi.SCHED_HALVE, sweeps_per_check=50, sweeps_per_exchange=10)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    for _ in range(4):
        for _ in range(5):
            eng.sweeps(10); ora.sweeps(10)
            eng.exchange(); ora.exchange()
        eng.check(); ora.check()
    assert_same(eng, ora, "after 200 sweeps", cursors=False)
    window_states_equal(eng, ora)


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
def test_batched_mixed_sizes_bit_exact(gpu, rng):
    """Several random-strain AT realisations of different L and one clean Ising
    system in one context (one launch per model/path class), with the
    visited-support flatness test: bit-exact with the oracle, which runs each
    system with its own streams."""
    print(f"case: batched mixed sizes rng={rng}", flush=True)
    systems, params = configs.config4(realisations=4, sizes=(4, 6, 8, 6), bins=64, windows=3, replicas=2,
                                      lnf_final=1e-3)
    systems.append(configs.ising_full_range(6, 4.0, 2, 3, seed=99))
    params = dict(params, rng_mode=rng, sweeps_per_check=40, sweeps_per_exchange=20, flatness=0.5)
    eng, ora = make_pair(systems, params)
    eng.adjust_to_range(); ora.adjust_to_range()
    for cycle in range(12):
        for _ in range(2):
            eng.sweeps(20); ora.sweeps(20)
            eng.exchange(); ora.exchange()
        eng.check(); ora.check()
    assert_same(eng, ora, "batched", cursors=(rng == abi.RNG_REPLAY))
    window_states_equal(eng, ora)
    assert eng.stats()["flat_events"] > 0


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
def test_ashkin_teller_full_spectrum_adjust_bit_exact(gpu, rng):
    """Full AT spectrum [-5N, 3N]: the plane-alternating range adjustment
    reaches the top windows; engine and oracle agree."""
    print(f"case: AT full spectrum adjust rng={rng}", flush=True)
    N = 100
    s = configs.at_system(10, -5.0 * N, 3.0 * N + 8.0, 8.0, 20, 2)
    p = configs.params(rng_mode=rng, adjust_mode=1, lnf_initial=0.05)
    eng, ora = make_pair([s], p)
    assert eng.windows[-1]["e_min"] > N   # above what sigma flips alone can reach
    eng.adjust_to_range(); ora.adjust_to_range()
    assert_same(eng, ora, "after adjust", cursors=(rng == abi.RNG_REPLAY))
    assert eng.energy(eng.num_walkers - 1) >= eng.windows[-1]["e_min"]
    eng.sweeps(50); ora.sweeps(50)
    assert_same(eng, ora, "after 50 sweeps", cursors=(rng == abi.RNG_REPLAY))


@pytest.mark.parametrize("kernel", [2, 3, 4])
@pytest.mark.parametrize("L", [16, 24, 32, 48, 64])
@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
def test_row_unrolled_kernel_bit_exact(gpu, model, L, kernel):
    """The row-unrolled kernel (sweep_kernel = 2; 3 with the next slot
    computed for both outcomes; 4 with the next ln g read ahead too) for every
    lattice size it is built for: narrow
    windows so that walkers keep proposing across the window edges (guard
    bins), a small ln f so that acceptances are frequent, and exchanges and
    checks in between; bit-exact with the oracle."""
    print(f"case: row-unrolled kernel model={model} L={L} kernel={kernel}", flush=True)
    N = L * L
    if model == abi.ISING2D:
        s = configs.ising_full_range(L, 8.0, 6, 2)
    else:
        s = configs.at_system(L, -5.0 * N, 3.0 * N + 8.0, 8.0, 6, 2)
    p = configs.params(sweep_kernel=kernel, lnf_initial=0.02, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.5,
                       schedule=abi.SCHED_HALVE, sweeps_per_check=20, sweeps_per_exchange=5,
                       adjust_mode=1 if model == abi.ASHKIN_TELLER2D else 0)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    for _ in range(3):
        for _ in range(4):
            eng.sweeps(5); ora.sweeps(5)
            eng.exchange(); ora.exchange()
        eng.check(); ora.check()
    assert_same(eng, ora, "after 60 sweeps", cursors=False)
    window_states_equal(eng, ora)


@pytest.mark.parametrize("field", ["gaussian", "dyadic"])
@pytest.mark.parametrize("L", [16, 24, 32, 48])
@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
def test_row_unrolled_field_walker_bit_exact(gpu, model, L, field):
    """The field-model kernel's row-unrolled walker (L in {16, 24, 32, 48})
    against the oracle: a Gaussian field (arbitrary FP64 energies, bins of 8),
    and a field of multiples of 1/2 with bins of width 3, where E' - min is
    often an exact multiple of the bin width, so that the reciprocal
    quotient lands on an integer and the walker divides (the exact
    fallback of its bin computation)."""
    print(f"case: row-unrolled field walker model={model} L={L} field={field}", flush=True)
    N = L * L
    if field == "gaussian":
        h = configs.gaussian_strain(L, 0.5, seed=100 + L)
        bs = 8.0
    else:
        h = np.round(configs.gaussian_strain(L, 1.0, seed=200 + L) * 2.0) / 2.0
        bs = 3.0
    span = float(np.ceil(np.abs(h).sum())) + 8.0
    if model == abi.ISING2D:
        s = configs.ising_system(L, -2.0 * N - span, 2.0 * N + span, bs, 6, 2, field=h)
    else:
        s = configs.at_system(L, -5.0 * N - span, 3.0 * N + span, bs, 6, 2, field=h)
    p = configs.params(lnf_initial=0.02, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.5,
                       schedule=abi.SCHED_HALVE, sweeps_per_check=20, sweeps_per_exchange=5,
                       adjust_mode=1 if model == abi.ASHKIN_TELLER2D else 0)
    eng, ora = make_pair([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    for _ in range(2):
        for _ in range(4):
            eng.sweeps(5); ora.sweeps(5)
            eng.exchange(); ora.exchange()
        eng.check(); ora.check()
    assert_same(eng, ora, "after 40 sweeps", cursors=False)
    window_states_equal(eng, ora)


def test_row_unrolled_kernel_needs_a_built_size(gpu):
    s = configs.ising_full_range(12, 8.0, 2, 2)
    with pytest.raises(abi.REWLError):
        Engine([s], configs.params(sweep_kernel=2))


def test_acceptance_bounds_device_equal_host(gpu):
    """The row-unrolled kernel's acceptance bounds B(k) (device FP64 log)
    equal the host's (glibc log), which tests/test_accept_bound.py pins
    against the reference's u < exp(-d) rule."""
    import ctypes as C
    lib = abi.load_engine()
    bad = C.c_double()
    abi.check(lib.rewl_selfcheck(1, C.byref(bad)))
    print(f"bounds differing from the host's: {bad.value:.0f} of 2^24", flush=True)
    assert bad.value == 0


def test_autotuner_picks_a_shape_that_runs(gpu):
    """The autotuner times every candidate launch shape of a small integer
    workload and returns the fastest; that shape runs the oracle's chain."""
    from rewl_simulator_amd import autotune
    systems = [configs.ising_full_range(16, 8.0, 2, 4)]
    p = configs.params(sweeps_per_check=20, sweeps_per_exchange=20)
    best, rates = autotune.autotune(systems, p, sweeps=8)
    assert best is not None and len(rates) >= 4
    assert all(r > 0 for r in rates.values())
    eng, ora = make_pair(systems, dict(p, **best))
    eng.adjust_to_range(); ora.adjust_to_range()
    eng.sweeps(10); ora.sweeps(10)
    assert_same(eng, ora, f"autotuned shape {best}", cursors=False)


def test_engine_autotune_creates_the_tuned_shape(gpu):
    """Engine(..., autotune=...) creates its context with the autotuner's
    choice (cached in a JSON file for the next process) and runs the oracle's
    chain with it."""
    import tempfile
    print("case: engine autotune", flush=True)
    systems = [configs.ising_full_range(16, 8.0, 2, 4)]
    p = configs.params(sweeps_per_check=20, sweeps_per_exchange=20)
    with tempfile.TemporaryDirectory() as d:
        cache = os.path.join(d, "shapes.json")
        eng = Engine(systems, p, autotune=cache)
        assert os.path.exists(cache)
        assert set(eng.tuned_shape) == {"walkers_per_group", "producers", "sweep_kernel"}
        assert all(eng.params[k] == v for k, v in eng.tuned_shape.items())
    ora = Oracle(systems, eng.params)
    eng.adjust_to_range(); ora.adjust_to_range()
    eng.sweeps(10); ora.sweeps(10)
    assert_same(eng, ora, f"engine autotuned {eng.tuned_shape}", cursors=False)
    eng.close(); ora.close()


def test_retired_systems_equal_their_own_runs(gpu):
    """rewl_params.retire = 1: in a batch, a system whose windows are all done
    stops sweeping, exchanging and checking, which is where a run of that
    system alone ends.  Each system of a mixed batch, run to completion, ends
    in the state of its own single-system run, bit for bit, and the batch is
    bit-exact with the oracle, which retires systems the same way."""
    print("case: retirement in a batch", flush=True)
    systems, params = configs.config4(realisations=4, sizes=(4, 6, 8, 6), bins=64, windows=3, replicas=2,
                                      lnf_final=2e-2)
    params = dict(params, sweeps_per_check=40, sweeps_per_exchange=20, flatness=0.5)
    assert params["retire"] == 1
    eng, ora = make_pair(systems, params)
    eng.adjust_to_range(); ora.adjust_to_range()
    checks = 0
    retired_seen = False
    while not eng.stats()["all_done"]:
        for _ in range(2):
            eng.sweeps(20); ora.sweeps(20)
            eng.exchange(); ora.exchange()
        eng.check(); ora.check()
        checks += 1
        done = [all(eng.window_info(k)["done"] for k in range(eng.num_windows) if eng.windows[k]["system"] == s)
                for s in range(len(systems))]
        retired_seen = retired_seen or (any(done) and not all(done))
        assert checks < 2000
    assert retired_seen   # some system finished while others still ran
    assert_same(eng, ora, "batch with retirement", cursors=False)
    window_states_equal(eng, ora)
    per = 3 * 2
    for si, s in enumerate(systems):
        with Engine([s], params) as alone:
            alone.adjust_to_range()
            while not alone.stats()["all_done"]:
                for _ in range(2):
                    alone.sweeps(20)
                    alone.exchange()
                alone.check()
            for j in range(per):
                g = si * per + j
                assert alone.energy(j) == eng.energy(g), (si, j)
                np.testing.assert_array_equal(alone.spins(j), eng.spins(g))
                np.testing.assert_array_equal(bits(alone.logdos(j)), bits(eng.logdos(g)))
    eng.close(); ora.close()
