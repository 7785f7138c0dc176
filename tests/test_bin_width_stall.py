"""Why the lowest windows of the L = 48 random-strain layout stall, on a
smaller lattice: bins wider than one excitation at the low-energy end.

Within a bin the Wang-Landau weight is a single number, so the walk inside a
bin follows the bare density of states, which across a bin of width w near
the ordered state grows by orders of magnitude (one flipped spin costs
12 +- 2 h_i here).  A walker leaves a bin downwards only from its lower edge,
where it rarely is once the bin spans two or more excitations.  The warm-start
layout (512 bins over [E(all-up), 0)) has bins of 10 at L = 32 and 22.5 at
L = 48, which is where the stall appears; this test shows the same switch at
L = 24 by changing only the bin width of one window over the bottom bins.
The oracle restates the engine's Philox-mode walk (the GPU runs are bit-exact
with it), so this is the algorithm's behaviour, not the device's.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rewl_simulator_amd import configs
from rewl_simulator_amd import warmstart as W


def _bottom_window(bin_width, checks=80):
    L, seed = 24, 924001
    h = configs.gaussian_strain(L, 0.5, seed)
    e0, _ = configs.rfat_range(L, h, K=0.5)
    s = configs.at_system(L, e0, e0 + 40 * bin_width, bin_width, 1, 4, K=0.5, field=h, seed=seed)
    o = Oracle([s], W.rfat_params(1e-6))
    o.adjust_to_range()          # every walker starts from all-up, in bin 0
    lowest = [10 ** 9] * 4
    for _ in range(checks):      # checks of 250 sweeps
        o.run(1)
        for g in range(4):
            lowest[g] = min(lowest[g], int((o.state(g)["E"] - e0) // bin_width))
    return o.window_state(0, 0)["lnf"], lowest


@pytest.mark.parametrize("bin_width", [10.0])
def test_bins_narrower_than_one_excitation_converge(bin_width):
    lnf, lowest = _bottom_window(bin_width)
    assert lnf < 1e-3, lnf
    assert max(lowest) <= 1, lowest    # every walker back at the bottom


@pytest.mark.parametrize("bin_width", [30.0])
def test_bins_of_two_excitations_stall(bin_width):
    lnf, lowest = _bottom_window(bin_width)
    # 2e4 sweeps later the window has not finished its first stage, and no
    # walker that left bin 0 at the start has come back near it
    assert lnf == 1.0, lnf
    assert min(lowest) >= 5, lowest
