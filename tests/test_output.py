"""The reference's result files (main.cpp:320-366): names, headers, column
order, delimiter and number format, read back the way the reference's own
plotting scripts read them (np.loadtxt, delimiter ",", comments "#")."""
import numpy as np
import pytest

from rewl_simulator_amd import abi, configs, output


def test_to_string_matches_std_to_string():
    assert output.to_string(16) == "16"
    assert output.to_string(1.0) == "1.000000"
    assert output.to_string(0.75) == "0.750000"
    assert output.to_string(1e-7) == "0.000000"
    assert output.to_string(-512.0) == "-512.000000"


def test_ising_strings():
    s = configs.ising_full_range(16, 4.0, 4, 2)
    model, base, hdr = output.system_strings(s)
    assert model == "Ising2d"
    assert base == "REWL_L-16_J-1.000000_h-0.000000.dat"
    lines = hdr.split("\n")
    assert lines[0] == "# REWL Ising2d on a Periodic Square Lattice"
    assert "#    L = 16" in lines and "#    N = 256" in lines and "#    num_DoF = 256" in lines
    assert "#    num_neighbors = 4" in lines
    assert "#    energy_min = -512.000000" in lines and "#    bin_size = 4.000000" in lines
    assert "#    num_bins = 257" in lines


def test_at_strings():
    s = configs.at_system(8, -320.0, 192.0, 1.0, 4, 2)
    model, base, hdr = output.system_strings(s)
    assert model == "Ashkin_Teller2d"
    assert base == "REWL_L-8_J-1.000000_K-0.500000.dat"
    assert "#    num_DoF = 128" in hdr.split("\n")
    h = configs.gaussian_strain(8, 0.5, 1)
    model, base, _ = output.system_strings(configs.at_system(8, -330.0, 0.0, 1.0, 4, 2, field=h))
    assert model == "Ashkin_Teller2d_RFAT_Baxter" and base.endswith("_h-0.000000.dat")


def test_rewl_header_and_file_header():
    r = output.rewl_header(8, 2, 0.75, 1000, 1e-6, 0.8)
    assert r.split("\n") == ["# REWL Parameters", "#    num_walkers = 8", "#    replicas_per_window = 2",
                             "#    window_overlap = 0.750000", "#    sweeps_per_check = 1000",
                             "#    final_increment = 0.000001", "#    flatness_criterion = 0.800000"]
    assert output.file_header("# A", "# B") == "# A\n#\n# B\n#\n"


def test_microcanonical_file_roundtrip(tmp_path):
    E = np.array([-512.0, -504.0, -500.0])
    lng = np.array([0.693, 6.2, 9.1])
    obs = np.array([[1.0, 1.0, 1.0, 10.0], [0.95, 0.91, 0.83, 20.0], [0.9, 0.81, 0.66, 30.0]])
    fn = output.write_microcanonical(str(tmp_path), "base.dat", "# H\n#\n", output.ISING_OBS, E, lng, obs, 256)
    assert fn.endswith("microcanonical_observables-base.dat.txt")
    text = open(fn).read()
    assert "#    1: Energy\n#    2: logDoS\n#    3: Magnetization\n#    4: Magnetization2\n" in text
    assert "#    6: Counts per Bin\n#\n" in text and "NUM OBS" not in text
    assert not text.endswith("\n")
    data = np.loadtxt(fn, delimiter=",", dtype="float64", comments="#")
    np.testing.assert_array_equal(data[:, 0], E / 256)
    np.testing.assert_array_equal(data[:, 1], lng / 256)
    np.testing.assert_array_equal(data[:, 2:], obs)
    assert "-2.000000000000000e+00,2.707031250000000e-03," in text


def test_nonlinear_closed_forms():
    T = np.array([1.0, 2.0])
    canon = np.tile([1.0, 1.0, 1.0], (2, 1))
    nl = output.nonlinear(abi.ISING2D, T, canon, 16)
    np.testing.assert_allclose(nl, [[0.0, 2.0 / 3.0], [0.0, 2.0 / 3.0]])
    c = np.tile([0.5, 0.3, 0.2, 0.5, 0.3, 0.2, 0.8, 0.7, 0.6, 0.4, 0.2, 0.1], (2, 1))
    nl = output.nonlinear(abi.ASHKIN_TELLER2D, T, c, 16)
    assert nl.shape == (2, 8)
    np.testing.assert_allclose(nl[:, 0], 16 * (0.3 - 0.25) / T)
    np.testing.assert_allclose(nl[:, 7], 2 * (1 - 0.5 * 0.6 / 0.49))


def test_nonlinear_correlation_length():
    """calculate_correlation_length (order_parameter_cumulants.hpp:53-59) over
    L, in the CORRELATION_LENGTHS column order of ising2d_observables.hpp:186-190
    and ashkin_teller2d_observables.hpp:285-297."""
    T = np.array([1.0, 2.0])
    L = 4
    # G0 = 5 Gq: xi/L = sqrt(4)/(2 pi); Gq = 0 gives DBL_MAX / L.
    canon = np.array([[0.5, 0.5, 0.3, 0.1], [0.5, 0.5, 0.3, 0.0]])
    nl = output.nonlinear(abi.ISING2D, T, canon, L * L, corr=True)
    assert nl.shape == (2, 3)
    np.testing.assert_allclose(nl[0, 2], 2.0 / (2 * np.pi), rtol=1e-15)
    assert nl[1, 2] == np.finfo(np.float64).max / L
    np.testing.assert_allclose(nl[:, :2], output.nonlinear(abi.ISING2D, T, canon[:, :3], L * L))
    # AT groups sigma, tau, phi, baxter of (m, m2, m4, G); output order sigma, tau, baxter, phi.
    g = {"s": [0.5, 0.4, 0.3, 0.1], "t": [0.6, 0.5, 0.4, 0.05], "p": [0.7, 0.9, 1.1, 0.3], "b": [0.2, 0.2, 0.1, 0.04]}
    c = np.tile(g["s"] + g["t"] + g["p"] + g["b"], (2, 1))
    nl = output.nonlinear(abi.ASHKIN_TELLER2D, T, c, L * L, corr=True)
    assert nl.shape == (2, 12)
    ref = output.nonlinear(abi.ASHKIN_TELLER2D, T, c[:, [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14]], L * L)
    np.testing.assert_allclose(nl[:, [0, 1, 3, 4, 6, 7, 9, 10]], ref)
    for col, (m2, gq) in zip([2, 5, 8, 11], [(0.4, 0.1), (0.5, 0.05), (0.2, 0.04), (0.9, 0.3)]):
        np.testing.assert_allclose(nl[:, col], np.sqrt((m2 - gq) / gq) / (2 * np.pi), rtol=1e-14)
    assert len(output.AT_NONLINEAR_CORR) == 12 and output.AT_NONLINEAR_CORR[11] == "Phi Correlation Length over L"
    assert len(output.AT_OBS_CORR) == 18 and output.AT_OBS_CORR[3] == "Sigma G(qmin)"
