"""The reference's result files, written from an engine run.

Three text files per system, in the reference's names, headers, column
order, delimiter and number format (main.cpp:320-366 with main_headers.hpp's
DELIMITER = ","); with observables = 2 in the reference's default
CORRELATION_LENGTHS layout (G(qmin) columns and correlation lengths), with
observables = 1 in its layout without them:

  microcanonical_observables-<base>.txt   write_microcanonical_observables.hpp:16-65
      E/N, ln g/N, the per-bin observable means and counts per bin
  self_averaged_observables-<base>.txt    write_self_averaged_observables.hpp:14-58
      T, F/N, U/N, U^2/N^2, S/N, C/N, the canonical observables, counts per bin
  nonlinear_observables-<base>.txt        write_nonlinear_observables.hpp
      T, susceptibilities and Binder cumulants (order_parameter_cumulants.hpp)

<base> and the header follow the model's parameter string
(ising2d_parameter_string.cpp, ashkin_teller2d_parameter_string.cpp) and
rewl_parameter_string.hpp; values are printed as std::to_string prints them.
Per-bin observables are the engine's pooled means over every walker holding
the bin (the reference averages the windows' running means by overlap count).
"""
import os

import numpy as np

from . import abi

DELIMITER = ","

ISING_OBS = ["Magnetization", "Magnetization2", "Magnetization4", "Counts per Bin", "NUM OBS"]
ISING_NONLINEAR = ["Susceptibility", "Binder Cumulant"]
AT_OBS = ["Sigma Mag", "Sigma Mag2", "Sigma Mag4", "Tau Mag", "Tau Mag2", "Tau Mag4", "Phi", "Phi2", "Phi4",
          "Baxter", "Baxter2", "Baxter4", "Counts per Bin", "NUM OBS"]
AT_NONLINEAR = ["Sigma Susceptibility", "Sigma Binder Cumulant", "Tau Susceptibility", "Tau Binder Cumulant",
                "Baxter Susceptibility", "Baxter Binder Cumulant", "Phi Susceptibility", "Phi Binder Cumulant"]
ENERGY_OBS = ["Free Energy", "Energy", "Energy2", "Entropy", "Specific Heat", "NUM OBS"]
ISING_OBS_CORR = ["Magnetization", "Magnetization2", "Magnetization4", "G(qmin)", "Counts per Bin", "NUM OBS"]
ISING_NONLINEAR_CORR = ["Susceptibility", "Binder Cumulant", "Correlation Length over L"]
AT_OBS_CORR = ["Sigma Mag", "Sigma Mag2", "Sigma Mag4", "Sigma G(qmin)", "Tau Mag", "Tau Mag2", "Tau Mag4",
               "Tau G(qmin)", "Phi", "Phi2", "Phi4", "Phi G(qmin)", "Baxter", "Baxter2", "Baxter4",
               "Baxter G(qmin)", "Counts per Bin", "NUM OBS"]
AT_NONLINEAR_CORR = ["Sigma Susceptibility", "Sigma Binder Cumulant", "Sigma Correlation Length over L",
                     "Tau Susceptibility", "Tau Binder Cumulant", "Tau Correlation Length over L",
                     "Baxter Susceptibility", "Baxter Binder Cumulant", "Baxter Correlation Length over L",
                     "Phi Susceptibility", "Phi Binder Cumulant", "Phi Correlation Length over L"]


def to_string(v):
    """std::to_string: integers as such, floating point as printf("%f")."""
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return "%f" % float(v)


def _sci(v):
    return "%.15e" % v   # std::scientific, setprecision(digits10 = 15)


def system_strings(system):
    """(model_name, file_name_base, system header) of one system description."""
    L = int(system["L"])
    N = L * L
    # num_bins as main.cpp:117 prints it: static_cast<size_t>((max - min) / bin)
    nb = int((system["e_max"] - system["e_min"]) / system["bin_size"])
    field = system.get("field") is not None or float(system.get("h", 0.0)) != 0.0
    if system["model"] == abi.ISING2D:
        model = "Ising2d" + ("_RFIM" if system.get("field") is not None else "")
        base = "REWL_L-%s_J-%s_h-%s.dat" % (to_string(L), to_string(float(system["J"])), to_string(float(system.get("h", 0.0))))
        ham = ["L = " + to_string(L), "N = " + to_string(N), "num_DoF = " + to_string(N),
               "J = " + to_string(float(system["J"])), "h = " + to_string(float(system.get("h", 0.0))),
               "num_neighbors = 4"]
    else:
        model = "Ashkin_Teller2d_RFAT_Baxter" if field else "Ashkin_Teller2d"
        base = "REWL_L-%s_J-%s_K-%s" % (to_string(L), to_string(float(system["J"])), to_string(float(system["K"])))
        if field:
            base += "_h-" + to_string(float(system.get("h", 0.0)))
        base += ".dat"
        ham = ["L = " + to_string(L), "N = " + to_string(N), "num_DoF = " + to_string(2 * N),
               "J = " + to_string(float(system["J"])), "K = " + to_string(float(system["K"]))]
        if field:
            ham.append("h = " + to_string(float(system.get("h", 0.0))))
        ham.append("num_neighbors = 4")
    header = "# REWL " + model + " on a Periodic Square Lattice\n#"
    header += "\n# Hamiltonian Parameters"
    for s in ham:
        header += "\n#    " + s
    header += "\n#"
    header += "\n# Wang Landau Parameters"
    header += "\n#    energy_min = " + to_string(float(system["e_min"]))
    header += "\n#    energy_max = " + to_string(float(system["e_max"]))
    header += "\n#    bin_size = " + to_string(float(system["bin_size"]))
    header += "\n#    num_bins = " + to_string(nb)
    return model, base, header


def rewl_header(num_walkers, replicas_per_window, window_overlap, sweeps_per_check, final_increment, flatness):
    """rewl_parameter_string.hpp:36-45 (window_overlap and flatness are floats there)."""
    h = "# REWL Parameters"
    h += "\n#    num_walkers = " + to_string(int(num_walkers))
    h += "\n#    replicas_per_window = " + to_string(int(replicas_per_window))
    h += "\n#    window_overlap = " + to_string(float(window_overlap))
    h += "\n#    sweeps_per_check = " + to_string(int(sweeps_per_check))
    h += "\n#    final_increment = " + to_string(float(final_increment))
    h += "\n#    flatness_criterion = " + to_string(float(flatness))
    return h


def file_header(system_header, rewl_hdr):
    """create_file_header (file_header.cpp:3-8)."""
    return system_header + "\n#\n" + rewl_hdr + "\n#\n"


def _names_block(first, names, title="# Intensive Observable Names by Column"):
    out = title
    counter = 0
    for n in first:
        counter += 1
        out += "\n#    %d: %s" % (counter, n)
    for n in names:
        if n != "NUM OBS":
            counter += 1
            out += "\n#    %d: %s" % (counter, n)
    return out + "\n#\n"


def write_microcanonical(path, base, header, obs_names, E, lng, obs, system_size):
    """obs[bin, ob] holds every observable column including counts per bin."""
    fn = os.path.join(path, "microcanonical_observables-" + base + ".txt")
    rows = []
    for b in range(len(E)):
        cols = [_sci(E[b] / system_size), _sci(lng[b] / system_size)] + [_sci(v) for v in obs[b]]
        rows.append(DELIMITER.join(cols))
    with open(fn, "w") as f:
        f.write(header + _names_block(["Energy", "logDoS"], obs_names) + "\n".join(rows))
    return fn


# the nonlinear file titles its column list differently
# (write_nonlinear_observables.hpp:29; write_self_averaged_observables.hpp:31)
NONLINEAR_TITLE = "# Intensive Nonlinear Observable Names by Column"


def write_temperature_table(path, prefix, base, header, names, T, values):
    fn = os.path.join(path, prefix + "-" + base + ".txt")
    rows = [DELIMITER.join([_sci(T[t])] + [_sci(v) for v in values[t]]) for t in range(len(T))]
    title = NONLINEAR_TITLE if prefix == "nonlinear_observables" else "# Intensive Observable Names by Column"
    with open(fn, "w") as f:
        f.write(header + _names_block(["Temperature"], names, title) + "\n".join(rows))
    return fn


def nonlinear(model, T, canon, N, corr=False):
    """Susceptibilities N (m2 - m^2)/T, Binder cumulants 1 - m4/(3 m2^2)
    (two-component for phi: 2 (1 - m4/(2 m2^2))) and, with correlators, the
    correlation length over L, (1/(2 pi)) sqrt((G0 - Gq)/Gq) with G0 = m2
    (order_parameter_cumulants.hpp), in the reference's order."""
    L = int(round(np.sqrt(N)))

    def susc(m2, m):
        return N * (m2 - m * m) / T

    def binder(m4, m2):
        return 1. - (m4 / (m2 * m2)) / 3.

    def xi(g0, gq):
        with np.errstate(invalid="ignore", divide="ignore"):
            v = (L / (2. * np.arccos(-1.))) * np.sqrt((g0 - gq) / gq)
        return np.where(gq == 0., np.finfo(np.float64).max, v) / L

    k = 4 if corr else 3
    c = canon
    if model == abi.ISING2D:
        out = [susc(c[:, 1], c[:, 0]), binder(c[:, 2], c[:, 1])]
        if corr:
            out.append(xi(c[:, 1], c[:, 3]))
        return np.stack(out, axis=1)
    sg, ta, ph, bx = 0, k, 2 * k, 3 * k
    phi_binder = 2. * (1. - 0.5 * (c[:, ph + 2] / (c[:, ph + 1] * c[:, ph + 1])))
    out = []
    for base in (sg, ta, bx):
        out += [susc(c[:, base + 1], c[:, base]), binder(c[:, base + 2], c[:, base + 1])]
        if corr:
            out.append(xi(c[:, base + 1], c[:, base + 3]))
    out += [susc(c[:, ph + 1], c[:, ph]), phi_binder]
    if corr:
        out.append(xi(c[:, ph + 1], c[:, ph + 3]))
    return np.stack(out, axis=1)


def write_results(eng, params, path, system=0, t_min=0.01, t_max=50.01, num_t=50000):
    """All three files for one system of a finished run (observables on,
    after rewl_finalize).  Returns the file names."""
    from .engine import thermo_observables, thermodynamics
    s = eng.systems[system]
    N = int(s["L"]) ** 2
    model, base, sys_hdr = system_strings(s)
    walkers = sum(1 for g in range(eng.num_walkers) if eng.windows[eng.walker_window[g]]["system"] == system)
    hdr = file_header(sys_hdr, rewl_header(walkers, s["replicas_per_window"], s.get("overlap", 0.75),
                                           params["sweeps_per_check"], params["lnf_final"], params["flatness"]))
    corr = params.get("observables", 0) == 2
    if model.startswith("Ising"):
        names, nl_names = (ISING_OBS_CORR, ISING_NONLINEAR_CORR) if corr else (ISING_OBS, ISING_NONLINEAR)
    else:
        names, nl_names = (AT_OBS_CORR, AT_NONLINEAR_CORR) if corr else (AT_OBS, AT_NONLINEAR)
    sps = 2.0 if s["model"] == abi.ISING2D else 4.0
    E, lng = eng.stitched(system, states_per_site=sps)
    Eg, means = eng.system_observables(system)
    counts = np.zeros(len(Eg))
    for w in eng.windows:
        if w["system"] != system:
            continue
        g0 = int(round((w["e_min"] - s["e_min"]) / s["bin_size"]))
        for r in range(s["replicas_per_window"]):
            g = w["first_walker"] + r
            counts[g0:g0 + w["num_bins"]] += eng.counts(g).astype(np.float64) + eng.visits(g)
    idx = np.searchsorted(Eg, E)
    ok = counts[idx] > 0
    E, lng = E[ok], lng[ok]
    obs = np.concatenate([means[idx][ok], counts[idx][ok, None]], axis=1)
    os.makedirs(path, exist_ok=True)
    files = [write_microcanonical(path, base, hdr, names, E, lng, obs, N)]
    T = t_min + (t_max - t_min) / num_t * np.arange(num_t)
    energy = thermodynamics(E, lng, N, t_min=t_min, t_max=t_max, num_t=num_t)
    canon = thermo_observables(E, lng, obs[:, :-1], t_min=t_min, t_max=t_max, num_t=num_t)
    counts_col = np.full((num_t, 1), obs[:, -1].mean())
    files.append(write_temperature_table(path, "self_averaged_observables", base, hdr, ENERGY_OBS + names, T,
                                         np.concatenate([energy, canon, counts_col], axis=1)))
    files.append(write_temperature_table(path, "nonlinear_observables", base, hdr, nl_names, T,
                                         nonlinear(s["model"], T, canon, N, corr)))
    return files
