"""Per-bin observables on the GPU against the oracle.  The engine keeps sums,
the oracle the reference's running means (update_observable_average,
ising2d_observables.hpp:134-145), so they agree to rounding: relative 1e-9.
The chains themselves stay bit-exact with observables on."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rewl_simulator_amd import abi, configs
from rewl_simulator_amd.engine import Engine

from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def _systems(model, field):
    L = 8
    N = L * L
    h = configs.gaussian_strain(L, 0.5, seed=11) if field else None
    if model == abi.ISING2D:
        return [configs.ising_system(L, -2.0 * N, 2.0 * N + 8.0, 8.0, 3, 2, field=h)]
    return [configs.at_system(L, -5.0 * N - 40, 3.0 * N + 40, 12.0, 3, 2, field=h)]


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
@pytest.mark.parametrize("field", [False, True])
@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
def test_observables_match_oracle(gpu, model, field, rng):
    print(f"case: observables model={model} field={field} rng={rng}", flush=True)
    systems = _systems(model, field)
    p = configs.params(rng_mode=rng, lnf_initial=0.05, observables=1, sweeps_per_exchange=20, adjust_mode=1)
    eng, ora = Engine(systems, p), Oracle(systems, p)
    eng.adjust_to_range(); ora.adjust_to_range()
    for _ in range(4):
        eng.sweeps(20); ora.sweeps(20)
        eng.exchange(); ora.exchange()
    eng.check(); ora.check()
    eng.sweeps(30); ora.sweeps(30)
    assert_same(eng, ora, "observables on", cursors=(rng == abi.RNG_REPLAY))
    for g in range(eng.num_walkers):
        sums = eng.observables(g)
        n = eng.counts(g).astype(np.float64) + eng.visits(g)
        means, counts = ora.observables(g)
        np.testing.assert_array_equal(n, counts, err_msg=f"counts per bin, walker {g}")
        seen = n > 0
        assert seen.sum() > 3
        np.testing.assert_allclose(sums[seen] / n[seen, None], means[seen], rtol=1e-9, atol=1e-14,
                                   err_msg=f"walker {g}")
        assert np.all(sums[~seen] == 0)
    eng.close(); ora.close()


def test_observables_off_refuses_query(gpu):
    systems = _systems(abi.ISING2D, False)
    with Engine(systems, configs.params()) as eng:
        with pytest.raises(abi.REWLError):
            eng.observables(0)


def test_canonical_observables_match_oracle(gpu):
    """The GPU canonical averages of per-bin observables against the oracle's
    restatement of the reference loop, on a converged-looking curve (exact
    L=16 Ising ln g) with smooth observables, 50 000 temperatures."""
    from oracle import oracle as O
    from rewl_simulator_amd import exact
    from rewl_simulator_amd.engine import thermo_observables
    E = -512.0 + 4.0 * np.arange(257)
    lng = exact.binned_lng(16, -512.0, 4.0, 257)
    ok = np.isfinite(lng)
    E, lng = E[ok], lng[ok]
    m = np.abs(E) / 512.0
    obs = np.stack([m, m * m, m ** 4], axis=1)
    g = thermo_observables(E, lng, obs)
    o = O.thermo_observables(E, lng, obs)
    np.testing.assert_allclose(g, o, rtol=1e-11, atol=1e-300)


def test_canonical_magnetisation_L4_end_to_end(gpu):
    """Observables on, a WL run of the 4x4 Ising model, pooled per-bin means,
    stitched ln g and the canonical averages on the GPU: <|m|>(T) and <m^2>(T)
    against the exact sums over all 2^16 states."""
    from rewl_simulator_amd.engine import thermo_observables
    L, N = 4, 16
    s = configs.ising_full_range(L, 4.0, 2, 2)
    p = configs.params(observables=1, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE,
                       lnf_final=1e-5, sweeps_per_check=500, sweeps_per_exchange=500, average_mode=1)
    with Engine([s], p) as eng:
        eng.adjust_to_range()
        assert eng.run(20000)
        eng.finalize()
        E, lng = eng.stitched(0, states_per_site=2.0)
        Eg, means = eng.system_observables(0)
    idx = np.searchsorted(Eg, E)
    obs = means[idx]
    ok = np.isfinite(obs[:, 0])
    T = 1.0 + 0.5 * np.arange(9)
    out = thermo_observables(E[ok], lng[ok], obs[ok], t_min=1.0, t_max=5.5, num_t=9)
    # exact per-level g and microcanonical means from all 2^16 states, then
    # the same trapezoid canonical average (end levels at half weight)
    bits = (np.arange(1 << N)[:, None] >> np.arange(N)) & 1
    sp = (1 - 2 * bits).reshape(-1, L, L)
    Ex = -((sp * np.roll(sp, 1, 1)).sum(axis=(1, 2)) + (sp * np.roll(sp, 1, 2)).sum(axis=(1, 2)))
    m = np.abs(sp.sum(axis=(1, 2))) / N
    lev = np.unique(Ex)
    g = np.array([(Ex == e).sum() for e in lev], dtype=np.float64)
    m1 = np.array([m[Ex == e].mean() for e in lev])
    m2 = np.array([(m[Ex == e] ** 2).mean() for e in lev])
    half = np.ones(len(lev)); half[0] = half[-1] = 0.5
    for i, t in enumerate(T):
        w = half * g * np.exp(-(lev - lev.min()) / t)
        np.testing.assert_allclose(out[i, 0], (w * m1).sum() / w.sum(), atol=0.02, err_msg=f"T = {t}")
        np.testing.assert_allclose(out[i, 1], (w * m2).sum() / w.sum(), atol=0.02, err_msg=f"T = {t}")


@pytest.mark.parametrize("mode", [1, 2])
def test_result_files_from_a_run(gpu, tmp_path, mode):
    """A finished run with observables writes the reference's three files
    (mode 2 in its CORRELATION_LENGTHS layout); they read back with the
    reference scripts' loader."""
    from rewl_simulator_amd import output
    s = configs.ising_full_range(8, 4.0, 2, 2)
    p = configs.params(observables=mode, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE,
                       lnf_final=1e-3, sweeps_per_check=200, sweeps_per_exchange=200, average_mode=1)
    with Engine([s], p) as eng:
        eng.adjust_to_range()
        assert eng.run(20000)
        eng.finalize()
        files = output.write_results(eng, p, str(tmp_path), t_min=0.5, t_max=5.5, num_t=50)
    names = [f.split("/")[-1] for f in files]
    assert names == ["microcanonical_observables-REWL_L-8_J-1.000000_h-0.000000.dat.txt",
                     "self_averaged_observables-REWL_L-8_J-1.000000_h-0.000000.dat.txt",
                     "nonlinear_observables-REWL_L-8_J-1.000000_h-0.000000.dat.txt"]
    micro = np.loadtxt(files[0], delimiter=",", comments="#")
    selfav = np.loadtxt(files[1], delimiter=",", comments="#")
    nl = np.loadtxt(files[2], delimiter=",", comments="#")
    k = 4 if mode == 1 else 5
    assert micro.shape[1] == 2 + k and selfav.shape == (50, 1 + 5 + k) and nl.shape == (50, 1 + k - 2)
    # <|m|> falls from ~1 at T = 0.5 toward small values at T = 5.5
    assert selfav[0, 6] > 0.95 and selfav[-1, 6] < 0.6
    assert np.all(np.isfinite(nl[:, :3]))
    if mode == 2:
        text = open(files[2]).read()
        assert "#    4: Correlation Length over L\n" in text
        assert "#    6: G(qmin)\n" in open(files[0]).read()
        # the ground state has G(qmin) = 0; xi/L is largest in the ordered phase
        assert abs(micro[np.argmin(micro[:, 0]), 5]) < 1e-12
        T = nl[:, 0]
        low, mid = np.argmin(np.abs(T - 1.0)), np.argmin(np.abs(T - 3.0))
        assert np.isfinite(nl[low, 3]) and np.isfinite(nl[mid, 3]) and nl[low, 3] > nl[mid, 3] > 0


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
def test_correlators_match_oracle(gpu, model, rng):
    """observables = 2: the Fourier correlators sampled at every 1000th update
    of a bin, in sweeps and after exchanges, against the oracle's restatement
    of the reference's correlator; the moments alongside, chains bit-exact."""
    print(f"case: correlators model={model} rng={rng}", flush=True)
    systems = _systems(model, field=True)
    p = configs.params(rng_mode=rng, lnf_initial=0.05, observables=2, sweeps_per_exchange=50, adjust_mode=1)
    eng, ora = Engine(systems, p), Oracle(systems, p)
    eng.adjust_to_range(); ora.adjust_to_range()
    for _ in range(6):
        eng.sweeps(50); ora.sweeps(50)
        eng.exchange(); ora.exchange()
    eng.check(); ora.check()
    assert_same(eng, ora, "correlators on", cursors=(rng == abi.RNG_REPLAY))
    corr = eng.correlator_slots(model)
    sampled = 0
    for g in range(eng.num_walkers):
        sums = eng.observables(g)
        n = eng.counts(g).astype(np.float64) + eng.visits(g)
        means, counts = ora.observables(g)
        np.testing.assert_array_equal(n, counts)
        seen = n > 0
        k = np.ceil(n[seen] / 1000.0)
        for c in range(sums.shape[1]):
            div = k if c in corr else n[seen]
            np.testing.assert_allclose(sums[seen, c] / div, means[seen, c], rtol=1e-9, atol=1e-13,
                                       err_msg=f"walker {g}, column {c}")
        sampled += int((n > 1000).sum())
    assert sampled > 0   # some bins took more than one correlator sample
    eng.close(); ora.close()


@pytest.mark.parametrize("rng", [abi.RNG_REPLAY, abi.RNG_PHILOX])
@pytest.mark.parametrize("field", [False, True])
@pytest.mark.parametrize("corr", [1, 2])
def test_densities_match_oracle(gpu, corr, field, rng):
    """densities = 1: the AT sigma-tau histograms of every update, in sweeps
    and after exchanges, bit-exact against the oracle's restatement of
    update_densities; every update lands in one cell."""
    print(f"case: densities observables={corr} field={field} rng={rng}", flush=True)
    systems = _systems(abi.ASHKIN_TELLER2D, field)
    p = configs.params(rng_mode=rng, lnf_initial=0.05, observables=corr, densities=1, sweeps_per_exchange=20,
                       adjust_mode=1)
    eng, ora = Engine(systems, p), Oracle(systems, p)
    eng.adjust_to_range(); ora.adjust_to_range()
    for _ in range(4):
        eng.sweeps(20); ora.sweeps(20)
        eng.exchange(); ora.exchange()
    eng.check(); ora.check()
    assert_same(eng, ora, "densities on", cursors=(rng == abi.RNG_REPLAY))
    for g in range(eng.num_walkers):
        d = eng.densities(g)
        np.testing.assert_array_equal(d, ora.densities(g), err_msg=f"walker {g}")
        np.testing.assert_array_equal(d.sum(axis=1), eng.counts(g).astype(np.uint64) + eng.visits(g))
    eng.close(); ora.close()


def test_density_orientation(gpu):
    """A walker held in the ground state sigma = +1, tau = -1 by a one-bin
    window counts every update in cell 0 * 32 + 31 (sigma is the column)."""
    L, N = 4, 16
    s = configs.at_system(L, -5.0 * N, -5.0 * N + 4.0, 4.0, 1, 1)
    p = configs.params(lnf_initial=0.5, observables=1, densities=1, exchange=0, adjust_mode=1)
    with Engine([s], p) as eng:
        eng.set_spins(0, np.tile(np.array([1, -1], np.int8), N))
        eng.adjust_to_range()
        eng.sweeps(3)
        d = eng.densities(0)
    assert d.shape == (1, 1024) and d[0, 31] == 3 * 2 * N and d[0].sum() == 3 * 2 * N


def test_densities_refused_without_observables_or_for_ising(gpu):
    with pytest.raises(abi.REWLError):
        Engine(_systems(abi.ASHKIN_TELLER2D, False), configs.params(densities=1))
    with pytest.raises(abi.REWLError):
        Engine(_systems(abi.ISING2D, False), configs.params(observables=1, densities=1))
    with Engine(_systems(abi.ASHKIN_TELLER2D, False), configs.params(observables=1)) as eng:
        with pytest.raises(abi.REWLError):
            eng.densities(0)


def test_density_plots_from_a_run(gpu, tmp_path):
    """The reference's density-plot files from an AT run: one file per bin of
    the concatenated windows, every value in [0, 1] (each walker's bin is
    normalised to a largest cell of 1 before the averages), the ground bin
    only in the four corner cells."""
    from rewl_simulator_amd import densities as D, output
    L, N = 8, 64
    s = configs.at_system(L, -5.0 * N, -2.0 * N, 4.0, 3, 2)
    p = configs.params(observables=1, densities=1, flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8,
                       schedule=abi.SCHED_HALVE, lnf_initial=0.1, sweeps_per_check=200, sweeps_per_exchange=200,
                       adjust_mode=1)
    with Engine([s], p) as eng:
        eng.adjust_to_range()
        for _ in range(10):
            eng.sweeps(200); eng.exchange(); eng.check()
        plots = D.system_densities(eng)
        nb_sys = int(round((s["e_max"] - s["e_min"]) / s["bin_size"]))
        first = eng.windows[0]
    assert plots.shape == (nb_sys, 1024) and first["e_min"] == s["e_min"]
    seen = ~np.isnan(plots).any(axis=1)
    assert seen.sum() > 10
    assert np.all((plots[seen] >= 0) & (plots[seen] <= 1)) and np.all(plots[seen].max(axis=1) > 0)
    ground = plots[0]
    if not np.isnan(ground).any():
        assert np.all(ground[np.setdiff1d(np.arange(1024), [0, 31, 992, 1023])] == 0)
    _, base, _ = output.system_strings(s)
    files = D.write_density_plots(str(tmp_path / "Density_Plots"), plots, base)
    assert len(files) == nb_sys
    assert np.loadtxt(files[0], comments="#").shape == (32, 32)


@pytest.mark.parametrize("obs,L", [(1, 16), (2, 16), (2, 48)])
@pytest.mark.parametrize("model", [abi.ISING2D, abi.ASHKIN_TELLER2D])
def test_observables_row_kernel_match_oracle(gpu, model, obs, L):
    """observables = 1 and 2 in Philox mode at L = 16 and 48 run on the
    row-unrolled kernel, which tracks the magnetisations in registers and,
    with 2, the correlator's Fourier sums (moved by each row's flips, sampled
    at every 1000th update of a bin): chains bit-exact, per-bin sums against
    the oracle's running means (rtol 1e-9)."""
    print(f"case: row-kernel observables model={model} observables={obs} L={L}", flush=True)
    N = L * L
    s = (configs.ising_system(L, -2.0 * N, 2.0 * N + 8.0, 8.0, 3, 3) if model == abi.ISING2D
         else configs.at_system(L, -5.0 * N, 3.0 * N + 12.0, 12.0, 3, 3))
    p = configs.params(rng_mode=abi.RNG_PHILOX, lnf_initial=0.05, observables=obs, sweeps_per_exchange=20,
                       adjust_mode=1)
    eng, ora = Engine([s], p), Oracle([s], p)
    eng.adjust_to_range(); ora.adjust_to_range()
    # enough updates that bins pass 1000 (several correlator samples) with obs = 2
    per = 20 if obs == 1 else (60 if L == 16 else 40)
    for _ in range(3 if obs == 1 else 10):
        eng.sweeps(per); ora.sweeps(per)
        eng.exchange(); ora.exchange()
    eng.check(); ora.check()
    eng.sweeps(20); ora.sweeps(20)
    assert_same(eng, ora, "row kernel, observables on", cursors=False)
    corr = eng.correlator_slots(model) if obs == 2 else []
    sampled = 0
    for g in range(eng.num_walkers):
        sums = eng.observables(g)
        n = eng.counts(g).astype(np.float64) + eng.visits(g)
        means, counts = ora.observables(g)
        np.testing.assert_array_equal(n, counts, err_msg=f"counts per bin, walker {g}")
        seen = n > 0
        k = np.ceil(n[seen] / 1000.0)
        for c in range(sums.shape[1]):
            div = k if c in corr else n[seen]
            np.testing.assert_allclose(sums[seen, c] / div, means[seen, c], rtol=1e-9, atol=1e-13,
                                       err_msg=f"walker {g}, column {c}")
        sampled += int((n > 1000).sum())
    assert obs == 1 or sampled > 0
    eng.close(); ora.close()
