/*
 * rewl_c.h -- C ABI of the MI355X Replica-Exchange Wang-Landau engine.
 *
 * The reference (meese-wj/REWL_Simulator) has no runtime plug-in boundary: its
 * model is chosen at compile time (src/Hamiltonians/model_hamiltonians.hpp:6-40)
 * and one MPI rank drives one walker (src/Simulation/rewl_simulation.hpp:69-108).
 * This header is the boundary that replaces that arrangement: one context owns
 * every walker of one or more REWL systems on one GPU, and the host talks to
 * the device only through the functions below.  Plain pointers and sizes; no
 * allocation crosses the boundary (callers own every output buffer); every
 * call returns an int status (0 = success, <0 = error class) and a
 * per-thread message is available from rewl_last_error().
 *
 * Entry points and the reference interface each one replaces:
 *   rewl_create            REWL_simulation ctor + glazier::construct_windows +
 *                          REWL_Walker ctor + import_DoFs/import_disorder
 *                          (rewl_simulation.hpp:69-108, glazier.hpp:69-115,
 *                           rewl_walker.hpp:118-139, main.cpp:201-208)
 *   rewl_adjust_to_range   REWL_Walker::adjust_state_to_range (rewl_walker.hpp:71-115)
 *   rewl_sweeps            REWL_Walker::wang_landau_walk -> Wang_Landau::
 *                          wang_landau_sweep/update (rewl_walker.hpp:141-158,
 *                          wang_landau.hpp:49-120)
 *   rewl_exchange          REWL_simulation::replica_exchange_update
 *                          (rewl_simulation.hpp:228-346)
 *   rewl_check             check_if_window_is_flat + average_and_redistribute_window
 *                          + reset_histogram + ln f schedule + done vote
 *                          (rewl_simulation.hpp:458-470, 361-454, 539-671)
 *   rewl_run               REWL_simulation::simulate main loop (rewl_simulation.hpp:473-691)
 *   rewl_finalize          final average_and_redistribute_window (rewl_simulation.hpp:687)
 *   rewl_get_X, rewl_set_X export_logdos / export_energy_bins / DoFs / current_energy
 *                          (rewl_histograms.hpp:101-106, rewl_walker.hpp:59-65, 187-205)
 *   rewl_stitch            concatenate_tables_multiple_overlap + normalize_logDoS
 *                          (concatenate_rewl_tables.hpp:167-305, array_shift.hpp:15-43)
 *   rewl_thermodynamics    Thermodynamics::calculate_thermodynamics (thermodynamics.hpp:156-236)
 *   rewl_random_field      generate_random_field with an explicit seed (random_fields.hpp:41-75,
 *                          random_number_generators.hpp:33-57)
 *   rewl_anneal            Simulated_Annealer::simulate_annealing (simulated_annealer.hpp:92-194)
 *   rewl_get_densities     Ashkin_Teller2d_Obs::density_histograms before transfer_and_normalize
 *                          (ashkin_teller2d_observables.hpp:113-138, ashkin_teller_densities.hpp:30-57)
 */
#ifndef REWL_C_H
#define REWL_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define REWL_OK                 0
#define REWL_ERR_INVALID       -1   /* bad configuration or argument            */
#define REWL_ERR_DEVICE        -2   /* HIP runtime error                         */
#define REWL_ERR_NONFINITE     -3   /* ln g became NaN/Inf                       */
#define REWL_ERR_ITER_CAP      -4   /* run stopped at max_checks, not converged  */
#define REWL_ERR_NO_DEVICE     -5   /* no GPU visible                            */
#define REWL_ERR_REPLAY        -6   /* replay stream exhausted                   */
#define REWL_ERR_RANGE         -7   /* a walker could not be brought into range  */
/* What an error leaves behind: REWL_ERR_INVALID, REWL_ERR_ITER_CAP and
 * REWL_ERR_RANGE leave the context usable (after REWL_ERR_RANGE, import other
 * spins or widen the windows and call rewl_adjust_to_range again; each call
 * starts from a clear range flag).  After REWL_ERR_REPLAY the replay streams
 * are spent and after REWL_ERR_DEVICE or REWL_ERR_NONFINITE the device state
 * is undefined: destroy the context. */

/* observables per bin (rewl_params.observables): Ising |m|, m^2, m^4; Ashkin-
 * Teller sigma |m|, m^2, m^4, tau |m|, m^2, m^4, phi, phi^2, phi^4, Baxter |m|,
 * m^2, m^4, in the reference's enum order without the correlator entries
 * (ising2d_observables.hpp:20-27, ashkin_teller2d_observables.hpp) */
#define REWL_NUM_OBS_ISING      3
#define REWL_NUM_OBS_AT        12
/* observables = 2 adds the Fourier correlator G(qmin) after each group of
 * three, the reference's CORRELATION_LENGTHS layout (Ising |m|, m^2, m^4, G;
 * AT sigma, tau, phi, Baxter each |x|, x^2, x^4, G), sampled at the 1st,
 * 1001st, ... update of a bin (fourier_correlator.hpp, ising2d_hamiltonian.hpp:
 * 247-252) */
#define REWL_NUM_OBS_ISING_CORR 4
#define REWL_NUM_OBS_AT_CORR   16
/* Ashkin-Teller sigma-tau densities (rewl_params.densities): per energy bin a
 * 32 x 32 count histogram of (sigma/N, tau/N) over [-1.05, 1.05]^2, cell
 * (tau axis) * 32 + (sigma axis), axis = (int)((x + 1.05) / (2.1 / 32))
 * (ashkin_teller_densities.hpp:11-34, ashkin_teller_densities_parameters.cxx) */
#define REWL_DENSITY_AXIS_BINS 32
#define REWL_DENSITY_CELLS     1024

/* ---- enumerations ----------------------------------------------------- */
enum rewl_model {
    REWL_ISING2D = 0,            /* Ising2d (ising2d_hamiltonian.hpp); per-site
                                    field = RFIM                                  */
    REWL_ASHKIN_TELLER2D = 1     /* Ashkin_Teller2d (ashkin_teller2d_hamiltonian.hpp);
                                    per-site field on sigma*tau = RFAT_BAXTER     */
};

enum rewl_flatness {
    REWL_FLAT_ALL_VISITED   = 0, /* ONE_OVER_T_ALGORITHM count_flatness:
                                    every bin visited (rewl_histograms.hpp:115-130)  */
    REWL_FLAT_RATIO         = 1, /* classic: (max-min)/min <= tol, float arithmetic
                                    (rewl_histograms.hpp:133-152)                    */
    REWL_FLAT_MIN_OVER_MEAN = 2, /* min >= tol * mean (the usual WL criterion)        */
    REWL_FLAT_MIN_OVER_MEAN_VISITED = 3  /* the same over the bins some replica of the
                                    window has visited: for continuous (random-field)
                                    spectra whose empty bins are not known in advance */
};

enum rewl_schedule {
    REWL_SCHED_HALVE      = 0,   /* ln f <- ln f / 2 on every flat event          */
    REWL_SCHED_ONE_OVER_T = 1    /* reference 1/t switch (rewl_simulation.hpp:557-567) */
};

enum rewl_rng {
    REWL_RNG_PHILOX  = 0,        /* counter-based Philox4x32-10, positional draws */
    REWL_RNG_REPLAY  = 1         /* per-walker std::mt19937_64 stream generated on
                                    the host and consumed on demand, exactly as the
                                    reference's random_number_generator
                                    (random_number_generators.hpp:14-59)          */
};

/* ---- configuration ------------------------------------------------------ */
typedef struct rewl_system_desc {
    int32_t model;               /* enum rewl_model                                  */
    int32_t L;                   /* periodic L x L square lattice, 2 <= L <= 64      */
    double  J;                   /* nearest-neighbour exchange                        */
    double  K;                   /* four-spin coupling (Ashkin-Teller only)           */
    double  h;                   /* uniform field (Ising: on sigma, AT: on sigma*tau) */
    const double* field;         /* L*L per-site fields, or NULL for the uniform h   */
    double  e_min;               /* global histogram range [e_min, e_max)            */
    double  e_max;
    double  bin_size;
    int32_t num_windows;
    int32_t replicas_per_window;
    double  overlap;             /* window overlap fraction (glazier EQUAL_WINDOWS)  */
    uint64_t seed;               /* Philox key / base seed of the replay streams     */
    const uint8_t* exclude_bins; /* per global bin: 1 = excluded from the flatness
                                    test (unreachable level), or NULL                */
    const uint64_t* walker_seeds;/* replay mode: one mt19937_64 seed per walker, or
                                    NULL for the reference's DIFFERENT_SEEDS=OFF
                                    choice (every walker seeded with 1)              */
} rewl_system_desc;

typedef struct rewl_params {
    int32_t flatness_mode;       /* enum rewl_flatness                               */
    double  flatness;            /* tolerance of the chosen flatness mode            */
    int32_t schedule;            /* enum rewl_schedule                               */
    double  lnf_initial;         /* reference: incrementer = 1 (rewl_walker.hpp:21)  */
    double  lnf_final;           /* done when ln f < lnf_final (rewl_simulation.hpp:594) */
    int32_t sweeps_per_check;    /* reference: 1000 (rewl_parameters.hpp:27)          */
    int32_t sweeps_per_exchange; /* reference: 1 (rewl_parameters.hpp:28)             */
    int32_t reduce_logdos;       /* REDUCE_LOGDOS: ln g -= ln g[0] after a flat event */
    int32_t rng_mode;            /* enum rewl_rng                                     */
    int32_t walkers_per_group;   /* walkers per wavefront (1..64); 0 = the largest
                                    count whose histograms and bound ring fit the
                                    LDS of one workgroup, per kernel               */
    int32_t exchange;            /* 0 disables replica exchange (independent walkers) */
    int32_t producers;           /* Philox fast path: wavefronts per workgroup that
                                    precompute the acceptance thresholds; 0 = auto,
                                    <0 = none (the walkers draw their own)          */
    int32_t adjust_mode;         /* range adjustment: 0 = the reference's scan, whose
                                    Ashkin-Teller flavour toggles twice per wrap so
                                    only sigma flips (enough for [E0, 0)); 1 = the
                                    same scan alternating sigma and tau passes, which
                                    reaches the upper half of the AT spectrum        */
    int32_t average_mode;        /* intra-window ln g average: 0 = weighted by each
                                    replica's counts, as rewl_simulation.hpp:393-397;
                                    1 = equal weights 1/R, the alternative left in
                                    rewl_simulation.hpp:401 (unbiased when replicas
                                    also exchange with other windows)               */
    int32_t sweep_kernel;        /* Philox integer path: 0 = auto (3 for L in
                                    {16,24,32,48,64}, else 1); 1 = row kernel;
                                    2 = row-unrolled; 3 = row-unrolled with the
                                    next slot computed for both outcomes; 4 = that
                                    and the next ln g read ahead (2-4 are an error
                                    for other L or producers < 0)                  */
    int32_t observables;         /* 1: per-bin observable sums of every update (each
                                    attempt and each post-exchange increment) at the
                                    walker's current bin: |m|, m^2, m^4 (Ising,
                                    ising2d_hamiltonian.hpp:234-256) or the 12 AT
                                    observables (ashkin_teller2d_hamiltonian.hpp:
                                    267-318); in Philox mode the row-unrolled
                                    kernel (integer energies, L in {16,24,32,48,64})
                                    or the field-model kernel keeps them, otherwise
                                    the generic one.
                                    2: also the Fourier correlators G(qmin) at every
                                    1000th update of a bin (REWL_NUM_OBS_*_CORR
                                    layout).  0: off                                */
    int32_t densities;           /* 1: Ashkin-Teller sigma-tau density histograms at
                                    every update (the AT_DENSITIES build,
                                    ashkin_teller2d_hamiltonian.hpp:291-293); needs
                                    observables >= 1 and Ashkin-Teller systems only  */
    int32_t retire;              /* 1: a system whose windows are all done stops
                                    (no sweeps, exchanges or checks; its walkers
                                    keep their state), as a reference run of that
                                    system alone would end there; 0: every system
                                    runs until all are done                         */
} rewl_params;

/* Per-window bookkeeping, one record per (system, window). */
typedef struct rewl_window_info {
    int32_t system;
    int32_t window;
    int32_t first_walker;        /* global index of replica 0                         */
    int32_t num_bins;
    double  e_min;               /* [e_min, e_max) of the window                      */
    double  e_max;
    double  bin_size;
    double  lnf;                 /* current ln f of the window                        */
    int64_t iteration;           /* flat events so far + 1                            */
    int64_t sweeps_in_iteration;
    int32_t done;                /* ln f < lnf_final                                  */
    int32_t one_over_t;          /* 1/t schedule engaged                              */
} rewl_window_info;

typedef struct rewl_stats {
    uint64_t sweeps;             /* sweeps per walker so far                          */
    uint64_t attempts;           /* spin-flip attempts, all walkers                   */
    uint64_t exchange_calls;
    uint64_t exchanges_accepted;
    uint64_t checks;
    uint64_t flat_events;
    int32_t  all_done;
    int32_t  num_walkers;
    double   kernel_ms_sweep;    /* accumulated sweep-kernel time (HIP events)        */
    double   kernel_ms_other;    /* exchange + check kernels                          */
    uint64_t sweep_launches;     /* sweep-kernel launches timed in kernel_ms_sweep    */
} rewl_stats;

typedef struct rewl_ctx rewl_ctx;

/* ---- library / device --------------------------------------------------- */
const char* rewl_last_error(void);
int rewl_device_count(int* count);
const char* rewl_version(void);

/* ---- lifetime ------------------------------------------------------------- */
int rewl_create(const rewl_system_desc* systems, int32_t num_systems,
                const rewl_params* params, rewl_ctx** out);
int rewl_destroy(rewl_ctx* ctx);

/* ---- simulation steps ----------------------------------------------------- */
/* rewl_sweeps (and rewl_run) return REWL_ERR_RANGE until rewl_adjust_to_range
 * has put every walker inside its window, and again after rewl_set_spins. */
int rewl_adjust_to_range(rewl_ctx* ctx);
int rewl_sweeps(rewl_ctx* ctx, int32_t num_sweeps);
int rewl_exchange(rewl_ctx* ctx);
int rewl_check(rewl_ctx* ctx);
/* One "check cycle" = sweeps_per_check sweeps with an exchange after every
 * sweeps_per_exchange sweeps, then rewl_check.  Runs until every window is
 * done or max_checks cycles have run (REWL_ERR_ITER_CAP). */
int rewl_run(rewl_ctx* ctx, int64_t max_checks);
int rewl_finalize(rewl_ctx* ctx);
/* No counterpart in the reference: a multicanonical refinement after the
 * Wang-Landau run.  The replicas of every window are averaged (rewl_finalize),
 * every window's ln f is set to 0, so that walkers sample with the fixed
 * weights 1/g_est(E), the visit and count histograms are cleared, and
 * `sweeps` sweeps run with a replica exchange after every sweeps_per_exchange
 * sweeps (when exchange is on); then every window's ln g(b) becomes
 * ln g(b) + ln H(b), H the window's histogram summed over its replicas (bins
 * no replica visited keep their ln g), for every replica.  The statistical
 * error of the result falls as 1 / sqrt(sweeps), without the saturation of
 * Wang-Landau's own estimate; stitch afterwards as after rewl_finalize. */
int rewl_refine(rewl_ctx* ctx, int64_t sweeps, int32_t sweeps_per_exchange);
/* rewl_refine in parts, for callers that report progress between them:
 * rewl_refine_begin (average, ln f = 0, clear), any number of
 * rewl_refine_sweeps (the exchange cadence counted across calls), then
 * rewl_refine_end (ln g += ln H).  begin + sweeps(a) + sweeps(b) + end is
 * rewl_refine(a + b), bit for bit. */
int rewl_refine_begin(rewl_ctx* ctx);
int rewl_refine_sweeps(rewl_ctx* ctx, int64_t sweeps, int32_t sweeps_per_exchange);
int rewl_refine_end(rewl_ctx* ctx);
int rewl_synchronize(rewl_ctx* ctx);

/* ---- queries ------------------------------------------------------------- */
int rewl_num_walkers(const rewl_ctx* ctx, int32_t* num_walkers);
int rewl_num_windows(const rewl_ctx* ctx, int32_t* num_windows);
int rewl_get_window_info(rewl_ctx* ctx, int32_t window_index, rewl_window_info* out);
int rewl_get_stats(rewl_ctx* ctx, rewl_stats* out);
int rewl_get_logdos(rewl_ctx* ctx, int32_t walker, double* out, int32_t n);
int rewl_get_counts(rewl_ctx* ctx, int32_t walker, uint32_t* out, int32_t n);
int rewl_get_visits(rewl_ctx* ctx, int32_t walker, uint64_t* out, int32_t n);
int rewl_get_spins(rewl_ctx* ctx, int32_t walker, int8_t* out, int32_t n);
int rewl_get_energy(rewl_ctx* ctx, int32_t walker, double* energy);
int rewl_get_flavor(rewl_ctx* ctx, int32_t walker, int32_t* flavor);
int rewl_get_rng_cursor(rewl_ctx* ctx, int32_t walker, uint64_t* cursor);
/* Observable sums of one walker (observables on): sums[b * nobs + o] = sum of
 * observable o over every update of the walker in bin b.  The reference keeps
 * running means (update_observable_average, ising2d_observables.hpp:134-145)
 * and counts_per_bin; here the mean is sums / (visits + counts) of that bin,
 * and for a correlator slot (observables = 2) sums / ceil((visits + counts) /
 * 1000), the number of samples.  nobs is REWL_NUM_OBS_ISING / _AT, or the
 * _CORR counts with observables = 2, for the walker's model. */
int rewl_get_observables(rewl_ctx* ctx, int32_t walker, double* sums, int32_t nbins, int32_t nobs);
/* Density histogram counts of one walker (densities on): out[b * 1024 + c] =
 * number of updates in bin b whose (sigma/N, tau/N) fell in cell c
 * (update_densities, ashkin_teller_densities.hpp:30-34).  The reference then
 * normalises each bin by its largest cell and averages the replicas
 * (transfer_and_normalize, density_averaging.cpp:9-28); that is host work. */
int rewl_get_densities(rewl_ctx* ctx, int32_t walker, uint32_t* out, int32_t nbins);

/* ---- state import (warm start, tests) ------------------------------------- */
int rewl_set_logdos(rewl_ctx* ctx, int32_t walker, const double* lng, int32_t n);
int rewl_set_window_lnf(rewl_ctx* ctx, int32_t window_index, double lnf);
int rewl_set_spins(rewl_ctx* ctx, int32_t walker, const int8_t* spins, int32_t n);
/* import_DoFs for every walker of one system (main.cpp:194-210: the master's
 * configuration is broadcast and each rank imports it): n = num_DoF values
 * +-1 in the reference's layout.  Energies are recomputed once for the whole
 * context; rewl_adjust_to_range must run before the next sweep. */
int rewl_set_system_spins(rewl_ctx* ctx, int32_t system, const int8_t* spins, int32_t n);

/* ---- results path ---------------------------------------------------------- */
/* Stitch the replica-0 ln g of every window of one system and normalise it to
 * states_per_site^N (full spectrum) or half of it when e_max == 0.  Writes at
 * most capacity entries; *n_out receives the curve length. */
int rewl_stitch(rewl_ctx* ctx, int32_t system, double states_per_site,
                double* energies, double* logdos, int32_t capacity, int32_t* n_out);
/* No counterpart in the reference: the same curve joined by the mean
 * difference of adjacent windows over every bin both have visited (replica 0
 * after rewl_finalize; a window takes over from the middle of that stretch),
 * on the system's global bin grid (bins no walker visited are NaN), then
 * normalised like rewl_stitch.  The reference's join takes the shift from one
 * bin pair, the first of the overlap when its derivatives match to an int. */
int rewl_stitch_overlap(rewl_ctx* ctx, int32_t system, double states_per_site,
                        double* energies, double* logdos, int32_t capacity, int32_t* n_out);
/* Canonical thermodynamics from (E, ln g) on the GPU, FP64, trapezoid weights:
 * out[t*5 + {0..4}] = free energy, U, U^2, S, C per site at T = Tmin + dT*t. */
int rewl_thermodynamics(const double* energies, const double* logdos, int32_t num_bins,
                        int32_t system_size, double t_min, double t_max, int32_t num_t,
                        double* out);
/* No counterpart in the reference: rewl_thermodynamics for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results, bit for bit;
 * its device staging is a static array of 2.75 MiB per GPU, held until the
 * process ends, so num_bins <= 16384 and num_t <= 65536 (REWL_ERR_INVALID
 * beyond).  Calls that overlap on one GPU share that array and race. */
int rewl_thermodynamics_noalloc(const double* energies, const double* logdos, int32_t num_bins,
                                int32_t system_size, double t_min, double t_max, int32_t num_t,
                                double* out);

/* Canonical averages of per-bin observables (the observable part of
 * Thermodynamics::calculate_thermodynamics, thermodynamics.hpp:156-236): on the
 * GPU, FP64, one thread per temperature, trapezoid weights (1/2 at the two end
 * bins); observables[b * num_obs + o] is the microcanonical mean of observable
 * o in bin b; out[t * num_obs + o] = <o> at T = t_min + dT t, dT = (t_max -
 * t_min)/num_t.  num_obs <= 16. */
int rewl_thermo_observables(const double* energies, const double* logdos, const double* observables,
                            int32_t num_bins, int32_t num_obs, double t_min, double t_max, int32_t num_t,
                            double* out);

/* ---- disorder ------------------------------------------------------------- */
enum rewl_disorder { REWL_DISORDER_UNIFORM = 0, REWL_DISORDER_GAUSSIAN = 1 };
/* generate_random_field (random_fields.hpp:41-75) with the clock seed replaced
 * by `seed`: one std::mt19937_64(seed); per site 20 discarded values and one
 * kept value of the chosen distribution: uniform 2 strength (-1 + 2 u), or
 * get_gaussian(0, strength) (random_number_generators.hpp:33-57: Box-Muller
 * on (u, v) redrawn while u <= DBL_EPSILON, then the cosine or the sine branch
 * by a third draw < 0.5).  Host only: needs no GPU. */
int rewl_random_field(uint64_t seed, int32_t distribution, double strength, int32_t num_sites, double* out);

/* ---- simulated annealing ------------------------------------------------- */
/* Simulated_Annealer (simulated_annealer.hpp:24-194, SIMULATED_ANNEALING
 * build): a ground-state search run before the REWL simulation, for systems
 * whose disorder makes the all-up state a poor start.  Defaults in brackets
 * are the reference's SA_Parameters (simulated_annealing_parameters.cxx). */
typedef struct rewl_anneal_params {
    double  initial_temperature; /* [25]                                           */
    double  final_temperature;   /* [0.25]; T_k = Ti (Tf / Ti)^(k / num_temperatures) */
    int32_t num_temperatures;    /* [250]; temperature steps k = 0 .. num_temperatures */
    int32_t block_size;          /* [1000] sweeps between equilibration checks     */
    double  stdev_tolerance;     /* [0.01] allowed relative change of the block
                                    energy variance                                 */
    double  frozen_criterion;    /* [0] a block stdev equal to it ends the step    */
    int32_t max_blocks;          /* blocks per temperature at most (no counterpart:
                                    the reference loops until equilibrated); >= 1  */
    int32_t chains;              /* independent chains per system (the reference
                                    runs one)                                       */
} rewl_anneal_params;
/* Anneal `chains` independent chains of every system on the GPU, one lane per
 * chain.  Each chain starts from random spins (randomize_dofs) and sweeps the
 * lattice (sigma pass then tau pass for AT, as the flavour of change_state
 * alternates) with Metropolis updates: accept if E' < E, else if
 * u < exp(beta (E - E')) (sa_update, simulated_annealer.hpp:66-86).  After
 * each block of block_size sweeps the block's mean and standard deviation
 * decide whether the temperature step is done (simulated_annealer.hpp:
 * 129-158, simulated_annealing_utilities.hpp).  Random numbers: Philox4x32-10
 * keyed by the system's seed, counter (attempt / 4, domain 4, chain), word
 * attempt mod 4, u = word 2^-32; the start spins from domain 5 with one word
 * per degree of freedom (u < 0.5: +1).
 * Outputs per (system, chain), index system * chains + chain: the lowest
 * energy met (min_energy_found), the final energy, the sweeps done; and, if
 * final_spins is not NULL, the final configuration in the reference's
 * spin_array layout (num_DoF int8 +-1 per chain, systems packed in order),
 * which main.cpp:56-76 hands to the REWL walkers as their start. */
int rewl_anneal(const rewl_system_desc* systems, int32_t num_systems, const rewl_anneal_params* params,
                double* min_energy, double* final_energy, uint64_t* sweeps, int8_t* final_spins);

/* ---- device self-check ----------------------------------------------------- */
/* No counterpart in the reference.  what = 0: largest absolute error of the
 * FP32 -ln(u) the row kernel screens acceptances with, over every 32-bit
 * Philox word (exhaustive); that kernel's acceptance band is 2^-12.
 * what = 1: number of 32-bit words, of 2^24 sampled (and 0, 1, 2^32-1,
 * 2^32-2), whose acceptance bound B(k) computed on the device differs from
 * the same computation on the host (glibc log).
 * what = 2: the device's LDS read bandwidth in GB/s, measured with 16-byte
 * conflict-free reads from every CU (eight workgroups of 256 threads per CU),
 * best of three timed launches: the peak of the memory tier the sweep kernels
 * keep their histograms in. */
int rewl_selfcheck(int32_t what, double* out);

#ifdef __cplusplus
}
#endif
#endif /* REWL_C_H */
