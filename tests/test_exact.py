"""Exact 2D Ising density of states: brute force, sum rules, closed forms."""
import itertools
import os
from collections import Counter

import numpy as np
import pytest

from rewl_simulator_amd import exact

HERE = os.path.dirname(os.path.abspath(__file__))


def _brute(L):
    N = L * L
    cnt = Counter()
    idx = np.arange(N)
    right = (idx % L + 1) % L + (idx // L) * L
    down = idx % L + ((idx // L + 1) % L) * L
    for s in range(1 << N):
        sp = 1 - 2 * ((s >> idx) & 1)
        E = -(sp * sp[right]).sum() - (sp * sp[down]).sum()
        cnt[int(E)] += 1
    return cnt


def test_brute_force_L4():
    cnt = _brute(4)
    E, g = exact.ising_dos(4)
    assert {int(e): v for e, v in zip(E, g) if v} == dict(cnt)
    # the published L=4 table
    assert dict(cnt) == {-32: 2, -24: 32, -20: 64, -16: 424, -12: 1728, -8: 6688, -4: 13568, 0: 20524,
                         4: 13568, 8: 6688, 12: 1728, 16: 424, 20: 64, 24: 32, 32: 2}


@pytest.mark.parametrize("L", [6, 8, 12, 16, 32])
def test_sum_rules_and_low_order(L):
    N = L * L
    E, g = exact.ising_dos(L)
    assert sum(g) == 2 ** N
    assert g == g[::-1]                       # E -> -E (bipartite lattice)
    assert all(v == 0 for v in g[1::2])       # odd numbers of broken bonds
    # b = 0, 2, 4, 6, 8 broken bonds: 2, 0, 2N (one flip), 4N (adjacent pair),
    # and b = 8: two separated flips + straight/bent triples + the 2x2 block
    assert g[0] == 2 and g[2] == 0 and g[4] == 2 * N and g[6] == 4 * N
    assert g[8] == 2 * (N * (N - 5) // 2 + 6 * N + N)


def test_L64_fixture_matches():
    path = os.path.join(HERE, "golden", "ising_L64_lng.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    d = np.load(path)
    assert d["E"].size == 2049 + 2048 - 2     # levels with g > 0 (E0+4, Emax-4 missing)
    assert d["lng"][0] == pytest.approx(np.log(2.0))
    # sum rule from the stored logarithms
    m = d["lng"].max()
    assert m + np.log(np.exp(d["lng"] - m).sum()) == pytest.approx(4096 * np.log(2.0), rel=1e-14)
