"""Ashkin-Teller sigma-tau density histograms (the reference's AT_DENSITIES
build, src/Hamiltonians/Ashkin_Teller2d/Density_Plots/) in the oracle and the
host post-processing: the cell index against exact rational arithmetic, the
oracle's per-bin histograms against the microcanonical distribution of the
3x3 model (brute force over all 2^18 states), and the normalisation, replica
average, window concatenation and file format against closed forms."""
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from rewl_simulator_amd import abi, configs, densities as D


def _exact_axis(M, N):
    """floor((M/N + 21/20) / (21/320)) in exact arithmetic."""
    v = (Fraction(M, N) + Fraction(21, 20)) / Fraction(21, 320)
    return v.numerator // v.denominator


def test_cell_index_matches_exact_arithmetic():
    # every magnetisation of L = 3..8 lattices, away from exact cell edges
    for L in range(3, 9):
        N = L * L
        M = np.arange(-N, N + 1, 2)
        got = D.axis_index(M / N)
        want = [_exact_axis(int(m), N) for m in M]
        np.testing.assert_array_equal(got, want, err_msg=f"L = {L}")
    assert D.axis_index(-1.0) == 0 and D.axis_index(0.0) == 16 and D.axis_index(1.0) == 31
    assert D.cell_index(1.0, 1.0) == 1023 and D.cell_index(1.0, -1.0) == 31 and D.cell_index(-1.0, 1.0) == 992


def _exact_cell_distribution_L3(J=1.0, K=0.5):
    """E -> probability of each of the 1024 cells over the states of energy E."""
    L, N = 3, 9
    idx = np.arange(1 << (2 * N), dtype=np.int64)
    bits = (idx[:, None] >> np.arange(2 * N)) & 1
    sg = (1 - 2 * bits[:, :N]).reshape(-1, L, L).astype(np.float64)
    ta = (1 - 2 * bits[:, N:]).reshape(-1, L, L).astype(np.float64)
    E = np.zeros(len(idx))
    for ax in (1, 2):
        s2, t2 = np.roll(sg, 1, ax), np.roll(ta, 1, ax)
        E -= (J * (sg * s2 + ta * t2) + K * sg * s2 * ta * t2).sum(axis=(1, 2))
    Ms = sg.sum(axis=(1, 2)).astype(int)
    Mt = ta.sum(axis=(1, 2)).astype(int)
    ax_of = {m: _exact_axis(m, N) for m in range(-N, N + 1, 2)}
    cell = np.array([ax_of[t] * 32 + ax_of[s] for s, t in zip(Ms, Mt)])
    out = {}
    for e in np.unique(E):
        m = E == e
        out[int(round(e))] = np.bincount(cell[m], minlength=1024) / m.sum()
    return out


def test_oracle_densities_match_microcanonical_L3():
    """Wang-Landau visits the states of one energy uniformly, so each bin's
    histogram converges to the distribution of (sigma, tau) cells over the
    states of that energy; every update lands in exactly one cell."""
    exact = _exact_cell_distribution_L3()
    L, N = 3, 9
    s = configs.at_system(L, -5.0 * N, 3.0 * N + 1.0, 1.0, 1, 1)
    o = O.Oracle([s], configs.params(lnf_initial=1e-3, observables=1, densities=1, exchange=0, adjust_mode=1,
                                     flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8,
                                     schedule=abi.SCHED_HALVE))
    o.adjust_to_range()
    o.sweeps(20000)
    dens = o.densities(0)
    _, counts = o.observables(0)
    np.testing.assert_array_equal(dens.sum(axis=1), counts)
    e_min, _, bs, nb = o.window_geometry(0, 0)
    checked = 0
    for b in range(nb):
        E = int(round(e_min + bs * b))
        if E not in exact:
            assert dens[b].sum() == 0
            continue
        assert np.all(dens[b][exact[E] == 0] == 0), f"E = {E}: a cell no state of this energy has"
        if counts[b] < 4000:
            continue
        np.testing.assert_allclose(dens[b] / counts[b], exact[E], atol=0.03, err_msg=f"E = {E}")
        checked += 1
    assert checked >= 15
    # the four ground states (sigma, tau uniform, +-1 each) sit in the corner cells
    assert dens[0].sum() > 0 and dens[0][[0, 31, 992, 1023]].sum() == dens[0].sum()
    o.close()


def test_oracle_density_orientation():
    """sigma indexes the column and tau the row (density_index,
    ashkin_teller_densities.hpp:23-27): a walker held in the ground state
    sigma = +1, tau = -1 (a one-bin window rejects every flip, the lowest
    excitation costs 8J + 4K) counts every update in cell 0 * 32 + 31."""
    L, N = 4, 16
    s = configs.at_system(L, -5.0 * N, -5.0 * N + 4.0, 4.0, 1, 1)
    o = O.Oracle([s], configs.params(lnf_initial=0.5, observables=1, densities=1, exchange=0, adjust_mode=1))
    o.set_spins(0, np.tile(np.array([1, -1], np.int8), N))   # (sigma, tau) per site
    o.adjust_to_range()
    o.sweeps(3)
    dens = o.densities(0)
    assert dens.shape == (1, 1024)
    assert dens[0, 31] == 3 * 2 * N and dens[0].sum() == 3 * 2 * N
    o.close()


def test_oracle_refuses_densities_for_ising():
    s = configs.ising_full_range(4, 4.0, 1, 1)
    with pytest.raises(Exception):
        O.Oracle([s], configs.params(observables=1, densities=1))


def test_normalise_and_window_average():
    c = np.zeros((3, 1024), np.uint32)
    c[0, [0, 5]] = [4, 2]
    c[1, 7] = 3
    n = D.normalise(c)
    assert n[0, 0] == 1.0 and n[0, 5] == 0.5 and n[1, 7] == 1.0 and n[0].sum() == 1.5
    assert np.all(np.isnan(n[2]))               # transfer_and_normalize: 0 / 0
    m = D.window_average([n, 0.5 * n, 0.25 * n])
    np.testing.assert_array_equal(m[:2], ((n + 0.5 * n) + 0.25 * n)[:2] * (1. / 3.))
    np.testing.assert_array_equal(D.window_average([n])[:2], n[:2])


class _FakeEngine:
    """Two windows of one system, two replicas each, bins 0..3 and 2..5."""

    def __init__(self, plots):
        self.systems = [dict(e_min=0.0, bin_size=1.0, replicas_per_window=2)]
        self.windows = [dict(system=0, first_walker=0, e_min=0.0), dict(system=0, first_walker=2, e_min=2.0)]
        self.plots = plots

    def densities(self, g):
        return self.plots[g]


def test_system_densities_running_overlap_average():
    rng = np.random.default_rng(5)
    plots = [rng.integers(1, 50, size=(4, 1024)).astype(np.uint32) for _ in range(4)]
    out = D.system_densities(_FakeEngine(plots))
    assert out.shape == (6, 1024)
    w0 = D.window_average([D.normalise(plots[0]), D.normalise(plots[1])])
    w1 = D.window_average([D.normalise(plots[2]), D.normalise(plots[3])])
    np.testing.assert_array_equal(out[:2], w0[:2])
    np.testing.assert_array_equal(out[2:4], 0.5 * ((1. * w0[2:4]) + w1[:2]))
    np.testing.assert_array_equal(out[4:], w1[2:])


def test_density_file_format(tmp_path):
    plot = np.full((2, 1024), 0.25)
    plot[0, 33] = 1.0
    plot[1, :] = np.nan
    files = D.write_density_plots(str(tmp_path), plot, "REWL_L-8_J-1.000000_K-0.500000.dat")
    assert [os.path.basename(f) for f in files] == ["density_plots_bin-0_REWL_L-8_J-1.000000_K-0.500000.dat.txt",
                                                    "density_plots_bin-1_REWL_L-8_J-1.000000_K-0.500000.dat.txt"]
    text = open(files[0]).read()
    lines = text.split("\n")
    assert lines[:10] == ["# Ashkin-Teller 2d Density Plot", "#", "# Density Plot Parameters:", "#",
                          "#     axis bins    = 32", "#     total bins   = 1024", "#     axis maximum = 1.050000",
                          "#     axis minimum = -1.050000", "#     binwidth     = 0.065625", "#"]
    rows = lines[10:]
    assert len(rows) == 32 and not text.endswith("\n")
    assert rows[1].split("  ")[1] == "1.000000000000000e+00" and rows[0].split("  ")[0] == "2.500000000000000e-01"
    data = np.loadtxt(files[0], comments="#")
    np.testing.assert_array_equal(data.ravel(), plot[0])
    assert open(files[1]).read().split("\n")[10].split("  ")[0] == "-nan"
