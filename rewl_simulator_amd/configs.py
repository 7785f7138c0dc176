"""System descriptions of the workloads this engine is measured and tested on.

Each builder returns ``(systems, params)``: a list of plain dicts (one per REWL
system, fields of ``rewl_system_desc``) and a dict of ``rewl_params`` fields.

Energies follow the reference's conventions: the Ising energy is
-J sum_<ij> s_i s_j - sum_i h_i s_i (ising2d_hamiltonian.hpp:162-210) and the
Ashkin-Teller energy -sum_<ij>[J(s s' + t t') + K s s' t t'] - sum_i h_i s_i t_i
(ashkin_teller2d_hamiltonian.hpp:169-217), both on an L x L periodic lattice.
"""
import numpy as np

from . import abi

# reference defaults (rewl_parameters.hpp:20-39, CMakeLists.txt:35-60)
REFERENCE_PARAMS = dict(
    flatness_mode=abi.FLAT_ALL_VISITED,
    flatness=0.3,
    schedule=abi.SCHED_ONE_OVER_T,
    lnf_initial=1.0,
    lnf_final=1e-7,
    sweeps_per_check=1000,
    sweeps_per_exchange=1,
    reduce_logdos=1,
    rng_mode=abi.RNG_PHILOX,
    walkers_per_group=0,
    exchange=1,
    producers=0,
    adjust_mode=0,
    average_mode=0,
    sweep_kernel=0,
    observables=0,
    densities=0,
    retire=0,
)

DEFAULT_SEED = 20240501


def params(**kw):
    p = dict(REFERENCE_PARAMS)
    p.update(kw)
    return p


def ising_system(L, e_min, e_max, bin_size, windows, replicas, overlap=0.75, J=1.0, h=0.0, field=None,
                 seed=DEFAULT_SEED, exclude=None, walker_seeds=None):
    return dict(model=abi.ISING2D, L=L, J=J, K=0.0, h=h, field=field, e_min=e_min, e_max=e_max,
                bin_size=bin_size, num_windows=windows, replicas_per_window=replicas, overlap=overlap,
                seed=seed, exclude_bins=exclude, walker_seeds=walker_seeds)


def at_system(L, e_min, e_max, bin_size, windows, replicas, overlap=0.75, J=1.0, K=0.5, h=0.0, field=None,
              seed=DEFAULT_SEED, exclude=None, walker_seeds=None):
    return dict(model=abi.ASHKIN_TELLER2D, L=L, J=J, K=K, h=h, field=field, e_min=e_min, e_max=e_max,
                bin_size=bin_size, num_windows=windows, replicas_per_window=replicas, overlap=overlap,
                seed=seed, exclude_bins=exclude, walker_seeds=walker_seeds)


def ising_unreachable_levels(L):
    """Ising levels on the periodic lattice that no configuration has:
    E0 + 4 and E_max - 4 (one broken bond is impossible; g(E0+4) = 0)."""
    N = L * L
    return (-2 * N + 4, 2 * N - 4)


def ising_full_range(L, bin_size, windows, replicas, overlap=0.75, **kw):
    """Full Ising spectrum [-2N, 2N] with the top level included: the global
    range is [-2N, 2N + bin) so the half-open reference windows hold E = 2N.
    Bins that contain only unreachable levels are excluded from flatness."""
    N = L * L
    e_min, e_max = -2.0 * N, 2.0 * N + bin_size
    nb = int(round((e_max - e_min) / bin_size))
    exclude = np.zeros(nb, dtype=np.uint8)
    lo = np.arange(nb) * bin_size + e_min
    for b in range(nb):
        levels = [e for e in range(int(lo[b]), int(lo[b] + bin_size)) if (e + 2 * N) % 4 == 0]
        if all(e in ising_unreachable_levels(L) for e in levels):
            exclude[b] = 1
    return ising_system(L, e_min, e_max, bin_size, windows, replicas, overlap, exclude=exclude, **kw)


def config1():
    """Plain REWL, 2D Ising, L=16, E in [-512, 512] step 4, 4 windows at 75%
    overlap, 2 walkers per window, flatness 0.8, ln f 1 -> 1e-6 by halving,
    replica exchange every 1000 sweeps, FP64 ln g."""
    s = ising_full_range(16, 4.0, 4, 2)
    p = params(flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE, lnf_final=1e-6,
               sweeps_per_check=1000, sweeps_per_exchange=1000)
    return [s], p


def config2(lnf_final=1e-8):
    """Plain REWL, 2D Ising, L=64 (N=4096, 2049 bins of width 8 over
    [-8192, 8192]), 32 windows at 75% overlap x 64 walkers, flatness 0.8,
    ln f to 1e-8; one lane per walker, bit-packed lattice."""
    s = ising_full_range(64, 8.0, 32, 64)
    p = params(flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE, lnf_final=lnf_final,
               sweeps_per_check=1000, sweeps_per_exchange=1000)
    return [s], p


def config3(lnf_final=1e-7):
    """Plain REWL, 2D Ashkin-Teller, L=32, J=1, K=0.5, 1024 bins of width 8 over
    [-5120, 3072] (top level included), 16 windows x 32 walkers, ln f to 1e-7."""
    N = 32 * 32
    s = at_system(32, -5.0 * N, 3.0 * N + 8.0, 8.0, 16, 32)
    p = params(flatness_mode=abi.FLAT_MIN_OVER_MEAN, flatness=0.8, schedule=abi.SCHED_HALVE, lnf_final=lnf_final,
               sweeps_per_check=1000, sweeps_per_exchange=1000, adjust_mode=1)
    return [s], p


def gaussian_strain(L, sigma, seed):
    """Per-site random strain h_i ~ N(0, sigma^2) from a realisation seed
    (numpy PCG64; the reference seeds its fields from the clock,
    random_fields.hpp:41-49, so no realisation of it can be reproduced)."""
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, sigma, size=L * L)


def reference_random_field(L, strength, seed, distribution=abi.DISORDER_GAUSSIAN):
    """The reference's generate_random_field (random_fields.hpp:41-75) for an
    explicit seed, computed by the engine library (rewl_random_field): with
    the seed the reference drew from the clock, this is its field."""
    import ctypes as C
    lib = abi.load_engine()
    out = np.empty(L * L, dtype=np.float64)
    abi.check(lib.rewl_random_field(C.c_uint64(seed), int(distribution), float(strength), L * L,
                                    out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def at_energy(L, sigma, tau, J=1.0, K=0.5, field=None):
    """Ashkin-Teller energy of a configuration (numpy bond sum)."""
    s = np.asarray(sigma, float).reshape(L, L)
    t = np.asarray(tau, float).reshape(L, L)
    E = 0.0
    for ax in (0, 1):
        ss = s * np.roll(s, -1, ax)
        tt = t * np.roll(t, -1, ax)
        E += float((-J * (ss + tt) - K * ss * tt).sum())
    if field is not None:
        E -= float((np.asarray(field).reshape(L, L) * s * t).sum())
    return E


def rfat_range(L, field, J=1.0, K=0.5):
    """[E(all-up), E(sigma = tau = checkerboard)] of a random-strain
    Ashkin-Teller realisation: the ground state for |h_i| < 4(J+K) and the
    state with every bond at its maximum."""
    N = L * L
    up = np.ones(N)
    chk = np.array([1.0 if (i % L + i // L) % 2 == 0 else -1.0 for i in range(N)])
    return at_energy(L, up, up, J, K, field), at_energy(L, chk, chk, J, K, field)


def config4(realisations=256, sizes=(16, 24, 32, 48), sigma=0.5, bins=512, windows=8, replicas=4,
            lnf_final=1e-6, base_seed=1000):
    """Batched random-strain Ashkin-Teller: H + -sum_i h_i s_i t_i with
    h_i ~ N(0, sigma^2) per realisation, sizes cycled over `sizes`, 8 windows x
    4 walkers each, `bins` uniform bins over [E(all-up), E(checkerboard)] with
    the top level included; all realisations in one context (one launch per
    sweep step); a realisation stops when all its windows are done.  Flatness
    is taken over the bins some replica has visited:
    the low end of a disordered spectrum has empty bins whose position depends
    on the realisation."""
    systems = []
    for r in range(realisations):
        L = sizes[r % len(sizes)]
        h = gaussian_strain(L, sigma, seed=base_seed + r)
        e0, e1 = rfat_range(L, h)
        bs = (e1 - e0) / (bins - 1)
        systems.append(at_system(L, e0, e0 + bs * bins, bs, windows, replicas, field=h, seed=base_seed + r))
    p = params(flatness_mode=abi.FLAT_MIN_OVER_MEAN_VISITED, flatness=0.8, schedule=abi.SCHED_HALVE,
               lnf_final=lnf_final, sweeps_per_check=1000, sweeps_per_exchange=1000, adjust_mode=1, retire=1)
    return systems, p
