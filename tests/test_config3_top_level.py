"""Why the Ashkin-Teller workload over the full spectrum (L = 32, top level
included, min >= 0.8 mean over every bin of a window) does not flatten: its
top bin holds only E = 3N, the four configurations with sigma and tau both
checkerboard (every bond at 2J - K), and the walkers of the top window do not
find them, while every lower window is flat long before.  Shown on the
oracle (the same dynamics as the GPU, bit for bit) with the L = 8 analogue:
4 windows x 2 replicas, bins of 8, 20 000 sweeps at ln f = 1 without checks."""
import numpy as np

from oracle.oracle import Oracle
from rewl_simulator_amd import abi, configs


def test_top_level_is_four_checkerboard_states():
    # brute force on L = 2 is degenerate (neighbours repeat); L = 4: count the
    # states at E = 3N by their definition, sigma and tau each one of the two
    # checkerboards, against the energy of every such configuration
    L, N = 4, 16
    J, K = 1.0, 0.5
    cb = np.array([[(x + y) % 2 for x in range(L)] for y in range(L)])
    top = 0
    for s in (cb, 1 - cb):
        for t in (cb, 1 - cb):
            sg, ta = 1 - 2 * s, 1 - 2 * t
            E = 0.0
            for y in range(L):
                for x in range(L):
                    for dx, dy in ((1, 0), (0, 1)):
                        u, v = (x + dx) % L, (y + dy) % L
                        E -= J * sg[y, x] * sg[v, u] + J * ta[y, x] * ta[v, u] + \
                            K * sg[y, x] * ta[y, x] * sg[v, u] * ta[v, u]
            top += E == 3 * N
    assert top == 4
    # and no configuration lies above: every bond is at most 2J - K
    assert max(-(J * a + J * b + K * a * b) for a in (-1, 1) for b in (-1, 1)) == 2 * J - K


def test_top_window_waits_on_the_top_level():
    L, N = 8, 64
    s = configs.at_system(L, -5.0 * N, 3.0 * N + 8.0, 8.0, 4, 2)
    p = configs.params(flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE,
                       lnf_initial=1.0, sweeps_per_check=1000, sweeps_per_exchange=1000, adjust_mode=1)
    o = Oracle([s], p)
    o.adjust_to_range()
    o.sweeps(20000)
    ratios = []
    for g in range(o.num_walkers):
        c = o.counts(g).astype(float)
        ratios.append((c.min() / c.mean(), c[-1]))
    o.close()
    # windows 0-2 (walkers 0-5): flat far beyond 0.8
    assert all(r >= 0.95 for r, _ in ratios[:6])
    # the top window's last bin (E = 3N) never visited by either replica
    assert all(last == 0 for _, last in ratios[6:])
