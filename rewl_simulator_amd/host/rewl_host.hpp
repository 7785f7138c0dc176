// rewl_host.hpp -- header-only C++17 host layer over the C ABI (rewl_c.h).
//
// It keeps the shape of the reference's compile-time plug-in surface so that a
// program written against it reads the same:
//   * model plug-ins Ising2d<data_t> and Ashkin_Teller2d<data_t>
//     (ising2d_hamiltonian.hpp:64-160, ashkin_teller2d_hamiltonian.hpp:76-167)
//     with their parameter blocks (ising2d_parameters.cxx:13-69,
//     ashkin_teller2d_parameters.cxx:13-70) and the random-field / random-strain
//     variants (RFIM, RFAT_BAXTER) as a per-site field;
//   * the selector aliases Hamiltonian_t / System_Parameters of
//     model_hamiltonians.hpp:6-40, chosen here by a template argument;
//   * REWL_Parameters (rewl_parameters.hpp:20-39);
//   * REWL_simulation<Hamiltonian> with simulate() and the results path of
//     main.cpp:201-365 (stitch, normalise, thermodynamics).
// Nothing here computes on the CPU: every step is a call into librewl_hip.so.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "rewl_c.h"

namespace rewl_host {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

inline void check(int rc) {
    if (rc != REWL_OK) throw Error(rc, std::string("rewl: ") + rewl_last_error());
}

// ---- model plug-ins ----------------------------------------------------------
// Parameters of the 2D Ising model: L x L periodic, H = -J sum s_i s_j - sum h_i s_i.
struct Ising2d_Parameters {
    int L = 24;                      // SYSTEM_SIZE_L
    double J = 1.0;                  // SYSTEM_EXCHANGE_J
    double h = 0.0;                  // SYSTEM_FIELD_H (uniform)
    std::vector<double> field;       // RFIM: one field per site (empty = uniform h)
    int N() const { return L * L; }
    int num_DoF() const { return N(); }
    double ground_state_degeneracy() const { return 2.0; }
    double states_per_site() const { return 2.0; }
    // Wang-Landau range of the reference: [-2JN, 0), bin 8J (ising2d_parameters.cxx:44-54)
    double energy_min() const { return -0.5 * J * 4.0 * N(); }
    double energy_max() const { return 0.0; }
    double energy_bin_size() const { return J * 4.0 * 2.0; }
};

// Ashkin-Teller: H = -sum [J(s s' + t t') + K s s' t t'] - sum h_i s_i t_i.
struct Ashkin_Teller2d_Parameters {
    int L = 24;
    double J = 1.0;                  // SYSTEM_EXCHANGE_J
    double K = 0.5;                  // SYSTEM_EXCHANGE_K
    double h = 0.0;                  // uniform Baxter field
    std::vector<double> field;       // RFAT_BAXTER: random strain per site
    int N() const { return L * L; }
    int num_DoF() const { return 2 * N(); }
    double ground_state_degeneracy() const { return 4.0; }
    double states_per_site() const { return 4.0; }
    // [-(2J+K) 2N, 0), bin 8(J+K) (ashkin_teller2d_parameters.cxx:35-41)
    double energy_min() const { return -0.5 * (2 * J + K) * 4.0 * N(); }
    double energy_max() const { return 0.0; }
    double energy_bin_size() const { return (J + K) * 4.0 * 2.0; }
};

template <typename data_t>
struct Ising2d {
    using Parameters = Ising2d_Parameters;
    static constexpr int32_t model_id = REWL_ISING2D;
    static void describe(const Parameters& P, rewl_system_desc& d) {
        d.model = model_id;
        d.L = P.L;
        d.J = P.J;
        d.K = 0.0;
        d.h = P.h;
        d.field = P.field.empty() ? nullptr : P.field.data();
    }
};

template <typename data_t>
struct Ashkin_Teller2d {
    using Parameters = Ashkin_Teller2d_Parameters;
    static constexpr int32_t model_id = REWL_ASHKIN_TELLER2D;
    static void describe(const Parameters& P, rewl_system_desc& d) {
        d.model = model_id;
        d.L = P.L;
        d.J = P.J;
        d.K = P.K;
        d.h = P.h;
        d.field = P.field.empty() ? nullptr : P.field.data();
    }
};

// ---- REWL parameters (rewl_parameters.hpp:20-39) ---------------------------------
struct REWL_Parameters {
    int num_windows = 8;
    int replicas_per_window = 1;
    double window_overlap = 0.75;
    int sweeps_per_check = 1000;
    int sweeps_per_exchange = 1;
    double final_increment = 1e-7;
    double flatness_criterion = 0.3;
    int flatness_mode = REWL_FLAT_ALL_VISITED;   // ONE_OVER_T_ALGORITHM default
    int schedule = REWL_SCHED_ONE_OVER_T;
    double initial_increment = 1.0;
    bool reduce_logdos = true;
    bool replay = false;                         // per-walker mt19937_64 streams
    uint64_t seed = 1;
    int adjust_mode = 0;                         // 1: alternate AT planes (full spectrum)
    int average_mode = 0;                        // 1: equal-weight replica average
    int observables = 0;                         // 1: per-bin observables (update_observables),
                                                 // 2: also the Fourier correlator G(qmin)
    bool densities = false;                      // AT sigma-tau density plots (AT_DENSITIES)
};

// generate_random_field (random_fields.hpp:41-75) with an explicit seed in
// place of the clock: the field the reference draws for that seed.
inline std::vector<double> random_field(uint64_t seed, double strength, int num_sites, bool gaussian = true) {
    std::vector<double> h((size_t)num_sites);
    check(rewl_random_field(seed, gaussian ? REWL_DISORDER_GAUSSIAN : REWL_DISORDER_UNIFORM, strength, num_sites,
                            h.data()));
    return h;
}

// SA_Parameters (simulated_annealing_parameters.cxx:9-20); max_blocks and
// chains have no counterpart there (the reference loops until a temperature
// step equilibrates, with one chain).
struct SA_Parameters {
    double initial_temperature = 25.;
    double final_temperature = 0.25;
    int num_iterations = 250;
    int block_size = 1000;
    double energy_stdev_tolerance = 0.01;
    double frozen_criterion = 0.;
    int max_blocks = 100;
    int chains = 1;
};

// The SIMULATED_ANNEALING start of main.cpp:56-76: anneal on the GPU and
// return the ground-state estimate (min_energy_found) for the bottom of the
// energy range, and the final configuration that every walker then imports.
// With several chains, the chain that met the lowest energy is taken.
struct Annealed {
    double ground_state_energy = 0.;
    std::vector<int8_t> spins;
};

template <class Hamiltonian_t>
Annealed simulated_annealing(const typename Hamiltonian_t::Parameters& sp, const SA_Parameters& sa, uint64_t seed) {
    rewl_system_desc d{};
    Hamiltonian_t::describe(sp, d);
    d.e_min = sp.energy_min();
    d.e_max = sp.energy_max();
    d.bin_size = sp.energy_bin_size();
    d.num_windows = 1;
    d.replicas_per_window = 1;
    d.seed = seed;
    rewl_anneal_params p{};
    p.initial_temperature = sa.initial_temperature;
    p.final_temperature = sa.final_temperature;
    p.num_temperatures = sa.num_iterations;
    p.block_size = sa.block_size;
    p.stdev_tolerance = sa.energy_stdev_tolerance;
    p.frozen_criterion = sa.frozen_criterion;
    p.max_blocks = sa.max_blocks;
    p.chains = sa.chains;
    const size_t C = (size_t)sa.chains, ndof = (size_t)sp.num_DoF();
    std::vector<double> emin(C), efin(C);
    std::vector<uint64_t> sweeps(C);
    std::vector<int8_t> spins(C * ndof);
    check(rewl_anneal(&d, 1, &p, emin.data(), efin.data(), sweeps.data(), spins.data()));
    const size_t best = (size_t)(std::min_element(emin.begin(), emin.end()) - emin.begin());
    Annealed out;
    out.ground_state_energy = emin[best];
    out.spins.assign(spins.begin() + (std::ptrdiff_t)(best * ndof), spins.begin() + (std::ptrdiff_t)((best + 1) * ndof));
    return out;
}

template <class Hamiltonian_t>
class REWL_simulation {
  public:
    using System_Parameters = typename Hamiltonian_t::Parameters;

    // energy range [emin, emax) with bins of bsize, as the reference's
    // REWL_simulation(all_energy_min, all_energy_max, all_energy_binsize)
    REWL_simulation(const System_Parameters& sp, const REWL_Parameters& rp, double emin, double emax, double bsize,
                    const std::vector<uint8_t>& exclude_bins = {})
        : sys_(sp), rp_(rp), exclude_(exclude_bins) {
        rewl_system_desc d{};
        Hamiltonian_t::describe(sys_, d);
        d.e_min = emin;
        d.e_max = emax;
        d.bin_size = bsize;
        d.num_windows = rp.num_windows;
        d.replicas_per_window = rp.replicas_per_window;
        d.overlap = rp.window_overlap;
        d.seed = rp.seed;
        d.exclude_bins = exclude_.empty() ? nullptr : exclude_.data();
        d.walker_seeds = nullptr;
        rewl_params p{};
        p.flatness_mode = rp.flatness_mode;
        p.flatness = rp.flatness_criterion;
        p.schedule = rp.schedule;
        p.lnf_initial = rp.initial_increment;
        p.lnf_final = rp.final_increment;
        p.sweeps_per_check = rp.sweeps_per_check;
        p.sweeps_per_exchange = rp.sweeps_per_exchange;
        p.reduce_logdos = rp.reduce_logdos ? 1 : 0;
        p.rng_mode = rp.replay ? REWL_RNG_REPLAY : REWL_RNG_PHILOX;
        p.walkers_per_group = 0;
        p.exchange = 1;
        p.producers = 0;
        p.adjust_mode = rp.adjust_mode;
        p.average_mode = rp.average_mode;
        p.observables = rp.observables;
        p.densities = rp.densities ? 1 : 0;
        check(rewl_create(&d, 1, &p, &ctx_));
    }

    // the reference's default range for the model
    REWL_simulation(const System_Parameters& sp, const REWL_Parameters& rp)
        : REWL_simulation(sp, rp, sp.energy_min(), sp.energy_max(), sp.energy_bin_size()) {}

    // import_DoFs (main.cpp:202): every walker starts from `spins` (num_DoF
    // values +-1 in the reference layout).  As in main.cpp:202-210 the walkers
    // are brought into their windows once, after the import (adjust_state_to_range,
    // or lazily by simulate), so each replay stream spends one start draw.
    void import_DoFs(const std::vector<int8_t>& spins) {
        check(rewl_set_system_spins(ctx_, 0, spins.data(), (int32_t)spins.size()));
        adjusted_ = false;
    }

    // REWL_Walker::adjust_state_to_range for every walker (main.cpp:210).
    void adjust_state_to_range() {
        check(rewl_adjust_to_range(ctx_));
        adjusted_ = true;
    }

    REWL_simulation(const REWL_simulation&) = delete;
    REWL_simulation& operator=(const REWL_simulation&) = delete;
    ~REWL_simulation() { rewl_destroy(ctx_); }

    // REWL_simulation::simulate: until every window is at the final increment.
    // Returns false when max_checks check cycles ran first.
    bool simulate(int64_t max_checks = INT64_MAX / 2) {
        if (!adjusted_) adjust_state_to_range();
        const int rc = rewl_run(ctx_, max_checks);
        if (rc == REWL_ERR_ITER_CAP) return false;
        check(rc);
        check(rewl_finalize(ctx_));
        return true;
    }

    // stitched and normalised (E, ln g), as main.cpp:309-325 produces them
    void results(std::vector<double>& energies, std::vector<double>& logdos) const {
        int32_t nw = 0, n = 0, cap = 0;
        check(rewl_num_windows(ctx_, &nw));
        for (int32_t k = 0; k < nw; ++k) {
            rewl_window_info wi{};
            check(rewl_get_window_info(ctx_, k, &wi));
            cap += wi.num_bins;
        }
        energies.assign(cap, 0.0);
        logdos.assign(cap, 0.0);
        check(rewl_stitch(ctx_, 0, sys_.states_per_site(), energies.data(), logdos.data(), cap, &n));
        energies.resize(n);
        logdos.resize(n);
    }

    // Thermodynamics::calculate_thermodynamics: rows of (F, U, U^2, S, C) per site
    std::vector<double> thermodynamics(double tmin = 0.01, double tmax = 50.01, int32_t nT = 50000) const {
        std::vector<double> E, lg, out((size_t)nT * 5);
        results(E, lg);
        check(rewl_thermodynamics(E.data(), lg.data(), (int32_t)E.size(), sys_.N(), tmin, tmax, nT, out.data()));
        return out;
    }

    // Per-bin observable means on the system's bin grid (observables on):
    // every walker's sums pooled over the windows and replicas holding the
    // bin, over their pooled update counts; means[b * nobs + o], NaN where no
    // walker has been.  A correlator column (observables = 2: column 3 of each
    // group of four) is sampled once per 1000 updates of a walker's bin, so
    // its sums are divided by the pooled sample counts, ceil(updates / 1000)
    // per walker.  The reference averages window means by overlap count.
    void observables(std::vector<double>& energies, std::vector<double>& means, int& nobs) const {
        int32_t nw = 0;
        check(rewl_num_windows(ctx_, &nw));
        rewl_window_info w0{};
        check(rewl_get_window_info(ctx_, 0, &w0));
        const bool ising = sys_.states_per_site() == 2.0;
        const bool corr = rp_.observables == 2;
        nobs = ising ? (corr ? REWL_NUM_OBS_ISING_CORR : REWL_NUM_OBS_ISING)
                     : (corr ? REWL_NUM_OBS_AT_CORR : REWL_NUM_OBS_AT);
        int nb = 0;
        std::vector<rewl_window_info> wins((size_t)nw);
        for (int32_t k = 0; k < nw; ++k) {
            check(rewl_get_window_info(ctx_, k, &wins[(size_t)k]));
            const int g0 = (int)std::llround((wins[(size_t)k].e_min - w0.e_min) / w0.bin_size);
            nb = std::max(nb, g0 + wins[(size_t)k].num_bins);
        }
        std::vector<double> sum((size_t)nb * nobs, 0.0), cnt((size_t)nb, 0.0), smp((size_t)nb, 0.0);
        for (int32_t k = 0; k < nw; ++k) {
            const rewl_window_info& w = wins[(size_t)k];
            const int g0 = (int)std::llround((w.e_min - w0.e_min) / w0.bin_size);
            std::vector<double> s((size_t)w.num_bins * nobs);
            std::vector<uint32_t> c((size_t)w.num_bins);
            std::vector<uint64_t> v((size_t)w.num_bins);
            for (int32_t r = 0; r < rp_.replicas_per_window; ++r) {
                const int32_t g = w.first_walker + r;
                check(rewl_get_observables(ctx_, g, s.data(), w.num_bins, nobs));
                check(rewl_get_counts(ctx_, g, c.data(), w.num_bins));
                check(rewl_get_visits(ctx_, g, v.data(), w.num_bins));
                for (int b = 0; b < w.num_bins; ++b) {
                    const double n = (double)c[(size_t)b] + (double)v[(size_t)b];
                    cnt[(size_t)(g0 + b)] += n;
                    smp[(size_t)(g0 + b)] += std::ceil(n / 1000.0);
                    for (int o = 0; o < nobs; ++o) sum[(size_t)(g0 + b) * nobs + o] += s[(size_t)b * nobs + o];
                }
            }
        }
        energies.resize((size_t)nb);
        means.assign((size_t)nb * nobs, std::nan(""));
        for (int b = 0; b < nb; ++b) {
            energies[(size_t)b] = w0.e_min + w0.bin_size * b;
            if (cnt[(size_t)b] > 0)
                for (int o = 0; o < nobs; ++o)
                    means[(size_t)b * nobs + o] =
                        sum[(size_t)b * nobs + o] / (corr && o % 4 == 3 ? smp[(size_t)b] : cnt[(size_t)b]);
        }
    }

    // Density plots of the system (densities on), one 1024-cell plot per bin
    // from the lowest window's first bin: each walker's bin divided by its
    // largest cell (transfer_and_normalize, ashkin_teller_densities.hpp:36-57;
    // 0/0 gives NaN there too), the replicas summed and scaled by 1/R
    // (density_averaging.cpp:9-28), and overlapping windows averaged with
    // equal weights in window order (concatenate_rewl_tables.hpp:261-263).
    void density_plots(std::vector<std::vector<double>>& plots) const {
        int32_t nw = 0;
        check(rewl_num_windows(ctx_, &nw));
        rewl_window_info w0{};
        check(rewl_get_window_info(ctx_, 0, &w0));
        const int R = rp_.replicas_per_window;
        plots.clear();
        std::vector<int> overlapping;
        for (int32_t k = 0; k < nw; ++k) {
            rewl_window_info w{};
            check(rewl_get_window_info(ctx_, k, &w));
            const size_t nb = (size_t)w.num_bins;
            std::vector<uint32_t> c(nb * REWL_DENSITY_CELLS);
            std::vector<double> avg(nb * REWL_DENSITY_CELLS, 0.0);
            for (int r = 0; r < R; ++r) {
                check(rewl_get_densities(ctx_, w.first_walker + r, c.data(), w.num_bins));
                for (size_t b = 0; b < nb; ++b) {
                    const uint32_t* cb = c.data() + b * REWL_DENSITY_CELLS;
                    uint32_t mx = 0;
                    for (int i = 0; i < REWL_DENSITY_CELLS; ++i) mx = std::max(mx, cb[i]);
                    for (int i = 0; i < REWL_DENSITY_CELLS; ++i)
                        avg[b * REWL_DENSITY_CELLS + i] += (double)cb[i] / (double)mx;
                }
            }
            if (R > 1)
                for (double& v : avg) v *= 1. / (double)R;
            const size_t g0 = (size_t)std::llround((w.e_min - w0.e_min) / w0.bin_size);
            for (size_t b = 0; b < nb; ++b) {
                const double* src = avg.data() + b * REWL_DENSITY_CELLS;
                if (g0 + b < plots.size()) {
                    const double n = (double)overlapping[g0 + b];
                    for (int i = 0; i < REWL_DENSITY_CELLS; ++i)
                        plots[g0 + b][(size_t)i] = (1. / (1. + n)) * ((1. * n) * plots[g0 + b][(size_t)i] + src[i]);
                    ++overlapping[g0 + b];
                } else {
                    plots.emplace_back(src, src + REWL_DENSITY_CELLS);
                    overlapping.push_back(1);
                }
            }
        }
    }

    rewl_stats stats() const {
        rewl_stats s{};
        check(rewl_get_stats(ctx_, &s));
        return s;
    }

    rewl_ctx* context() const { return ctx_; }

  private:
    System_Parameters sys_;
    REWL_Parameters rp_;
    std::vector<uint8_t> exclude_;
    rewl_ctx* ctx_ = nullptr;
    bool adjusted_ = false;   // walkers inside their windows (adjust_state_to_range ran after the last import)
};

}  // namespace rewl_host
