"""Ashkin-Teller sigma-tau density plots: the host side of the reference's
AT_DENSITIES build (src/Hamiltonians/Ashkin_Teller2d/Density_Plots/).

The engine counts, per walker and energy bin, the updates whose (sigma/N,
tau/N) fall in each cell of a 32 x 32 grid over [-1.05, 1.05]^2
(rewl_params.densities, update_densities at ashkin_teller2d_hamiltonian.hpp:
291-293).  What follows the run is here, in the reference's order:

  normalise          transfer_and_normalize (ashkin_teller_densities.hpp:36-57):
                     each bin divided by its largest cell (0/0 = NaN, as there)
  window_average     average_density_in_window (density_averaging.cpp:9-28):
                     replica sum times 1/R
  system_densities   the AT_DENSITIES part of concatenate_tables_multiple_overlap
                     (concatenate_rewl_tables.hpp:207-208, 261-263, 298-300):
                     running equal-weight average over overlapping windows
  write_density_plots  write_density_plots.cpp:14-45 with the header of
                     ashkin_teller_densities_parameter_string.cpp:6-18
"""
import os

import numpy as np

from . import abi

AXIS_BINS = abi.DENSITY_AXIS_BINS
CELLS = abi.DENSITY_CELLS
SIGMA_TAU_MAX = 1.05
SIGMA_TAU_MIN = -SIGMA_TAU_MAX
BINWIDTH = (SIGMA_TAU_MAX - SIGMA_TAU_MIN) / float(AXIS_BINS)


def axis_index(x):
    """density_axis_index (ashkin_teller_densities.hpp:17-21): the truncated
    quotient (x - min) / binwidth as uint32."""
    return ((np.asarray(x, dtype=np.float64) - SIGMA_TAU_MIN) / BINWIDTH).astype(np.uint32)


def cell_index(sigma, tau):
    """density_index (ashkin_teller_densities.hpp:23-27)."""
    return axis_index(tau).astype(np.int64) * AXIS_BINS + axis_index(sigma).astype(np.int64)


def normalise(counts):
    """counts [bin, 1024] -> float64 [bin, 1024] over each bin's largest cell."""
    c = np.asarray(counts)
    mx = c.max(axis=1, keepdims=True).astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        return c.astype(np.float64) / mx


def window_average(normalised):
    """Replica average of one window's normalised densities, list over replicas."""
    if len(normalised) == 1:
        return normalised[0].copy()
    s = normalised[0].copy()
    for x in normalised[1:]:
        s = s + x
    return s * (1. / float(len(normalised)))


def system_densities(eng, system=0):
    """Concatenated density plots [bin, 1024] of one system (densities on),
    on the system's bins from the lowest window's first bin to the highest
    window's last."""
    s = eng.systems[system]
    final, overlapping = [], []
    for w in eng.windows:
        if w["system"] != system:
            continue
        R = s["replicas_per_window"]
        win = window_average([normalise(eng.densities(w["first_walker"] + r)) for r in range(R)])
        if not final:
            g_first = int(round((w["e_min"] - s["e_min"]) / s["bin_size"]))
            final = [win[b] for b in range(len(win))]
            overlapping = [1] * len(win)
            continue
        start = int(round((w["e_min"] - s["e_min"]) / s["bin_size"])) - g_first
        if start < 0 or start > len(final):
            raise ValueError("windows are not in ascending, overlapping order")
        for j in range(len(win)):
            i = start + j
            if i < len(final):
                n = overlapping[i]
                final[i] = (1. / (1. + n)) * ((1. * n) * final[i] + win[j])
                overlapping[i] += 1
            else:
                final.append(win[j])
                overlapping.append(1)
    return np.array(final)


def density_header():
    """AT_Density_Parameters::Strings::header (std::to_string values)."""
    h = "# Ashkin-Teller 2d Density Plot\n"
    h += "#\n"
    h += "# Density Plot Parameters:\n"
    h += "#\n"
    h += "#     axis bins    = %d\n" % AXIS_BINS
    h += "#     total bins   = %d\n" % CELLS
    h += "#     axis maximum = %f\n" % SIGMA_TAU_MAX
    h += "#     axis minimum = %f\n" % SIGMA_TAU_MIN
    h += "#     binwidth     = %f\n" % BINWIDTH
    h += "#"
    return h


def _value(v):
    # iostream scientific with digits10 = 15; 0/0 prints as glibc prints the x86 default NaN
    return "-nan" if np.isnan(v) else "%.15e" % v


def write_density_plots(path, float_data, file_string, header=None):
    """One file per energy bin, density_plots_bin-<b>_<file_string>.txt: the
    header, then 32 rows of 32 values separated by two spaces."""
    header = density_header() if header is None else header
    os.makedirs(path, exist_ok=True)
    files = []
    for b, plot in enumerate(float_data):
        fn = os.path.join(path, "density_plots_bin-%d_%s.txt" % (b, file_string))
        rows = ["  ".join(_value(v) for v in plot[r * AXIS_BINS:(r + 1) * AXIS_BINS]) for r in range(AXIS_BINS)]
        with open(fn, "w") as f:
            f.write(header + "\n" + "\n".join(rows))
        files.append(fn)
    return files
