"""Writes tests/golden/ising_L{16,32,64}_lng.npz: exact ln g(E) of the
periodic 2D Ising model (J = 1) over the levels with g > 0, from the exact
integer density of states of rewl_simulator_amd.exact (Kaufman/Beale
polynomial, computed modulo primes and lifted by CRT).

    python tests/golden/make_exact_fixtures.py [L ...]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from rewl_simulator_amd import exact  # noqa: E402


def main(argv):
    sizes = [int(a) for a in argv[1:]] or [16, 32, 64]
    for L in sizes:
        E, lng = exact.ising_lng(L)
        path = os.path.join(HERE, f"ising_L{L}_lng.npz")
        np.savez_compressed(path, E=E.astype(np.int64), lng=lng, L=np.int64(L))
        print(path, E.size)


if __name__ == "__main__":
    main(sys.argv)
