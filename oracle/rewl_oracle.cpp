/*
 * rewl_oracle.cpp -- CPU restatement of the reference REWL algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke test
 * and the CPU-baseline leg of bench.py load it to produce the values the HIP
 * engine must reproduce.  The product never links, loads or calls it.
 *
 * It restates, in plain single-threaded C++ over FP64 spins exactly as the
 * reference stores them, the reference's per-walker Wang-Landau update and
 * sweep, window construction, range adjustment, replica exchange, flatness
 * test, count-weighted window averaging, ln f schedule, stitching,
 * normalisation and canonical thermodynamics.  Every FP64 expression keeps the
 * reference's operation order so that the engine's integer fast path and its
 * FP64 field path can be compared bit for bit.  Build with -ffp-contract=off.
 *
 * Two random streams are supported per walker:
 *   replay  std::mt19937_64 + std::uniform_real_distribution<double>(0,1) and
 *           std::shuffle from this toolchain's libstdc++ -- the very library the
 *           reference calls (random_number_generators.hpp:14-59,
 *           rewl_simulation.hpp:250), consumed on demand in reference order;
 *   philox  Philox4x32-10 (Salmon et al., SC'11), positional: sweep draw for
 *           attempt p of sweep n of walker w is word (p & 3) of
 *           Philox(ctr = {lo(n*B + p/4), hi(n*B + p/4), 0, w}, key = seed),
 *           B = ceil(num_DoF/4), u = word * 2^-32; the range-adjust draw is
 *           Philox(ctr={0,0,1,w}); exchange draws are sequential 64-bit words
 *           Philox(ctr={lo(c),hi(c),2,w}) with a per-walker cursor c.
 *
 * With rewl_params.observables it also keeps the reference's per-bin
 * observable running means and counts per bin (update_observables,
 * update_observable_average), the Fourier correlator with observables = 2,
 * with rewl_params.densities the AT sigma-tau density counts
 * (update_densities), and restates the observable part of the canonical
 * thermodynamics.
 *
 * Parity pins (tests/test_oracle.py, tests/test_observables.py,
 * tests/test_densities.py):
 * mt19937_64 standard KAT, Philox Random123 KATs, neighbour-table fixture of
 * the reference's own tester, closed-form energies, brute-force L=4 density
 * of states, exact Beale DoS, count conservation at the range test's slack,
 * microcanonical observables against brute force over all states (L=4
 * Ising, L=3 Ashkin-Teller), G(qmin) against a brute-force structure factor,
 * density cells against exact rational cell indices and the brute-force
 * microcanonical (sigma, tau) distribution of the 3x3 AT model, a pinned
 * ground state for the cell orientation, two- and three-level canonical
 * closed forms.
 */
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../include/rewl_c.h"

namespace {

thread_local std::string g_err;
int fail(int code, const std::string& msg) { g_err = msg; return code; }

/* ---------------- Philox4x32-10 (Salmon, Moraes, Dror, Shaw 2011) -------- */
struct P4 { uint32_t v[4]; };
inline P4 philox4x32_10(P4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += W0; k1 += W1; }
        uint64_t p0 = (uint64_t)M0 * c.v[0];
        uint64_t p1 = (uint64_t)M1 * c.v[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        P4 o;
        o.v[0] = hi1 ^ c.v[1] ^ k0;
        o.v[1] = lo1;
        o.v[2] = hi0 ^ c.v[3] ^ k1;
        o.v[3] = lo0;
        c = o;
    }
    return c;
}

/* libstdc++ generate_canonical<double,53> over a 64-bit engine: one call,
 * double(x) * 2^-64, pulled below 1 when the conversion rounds up. */
inline double canonical_from_u64(uint64_t x) {
    double u = (double)x * 0x1p-64;
    if (u >= 1.0) u = std::nextafter(1.0, 0.0);
    return u;
}

enum { DOMAIN_SWEEP = 0, DOMAIN_INIT = 1, DOMAIN_EXCH = 2 };

/* URBG over a walker's Philox exchange stream, so that std::shuffle runs the
 * same algorithm on it as on mt19937_64. */
struct PhiloxExchangeUrbg {
    using result_type = uint64_t;
    uint32_t k0, k1, walker;
    uint64_t* cursor;
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return ~(uint64_t)0; }
    result_type operator()() {
        uint64_t c = (*cursor)++;
        P4 o = philox4x32_10(P4{{(uint32_t)c, (uint32_t)(c >> 32), (uint32_t)DOMAIN_EXCH, walker}}, k0, k1);
        return (uint64_t)o.v[0] | ((uint64_t)o.v[1] << 32);
    }
};

/* ---------------- histogram_index (histogram_index.hpp:8-31) ------------- */
struct HistIndex {
    double min, max, bs;
    size_t nbins;
    /* (size_t)((v - min) / bs) where the reference defines it.  The +0.5
     * range test below truncates toward zero, so with min in (-0.5, 0.5) it
     * admits v down to -1.5: for bins narrower than that slack the quotient
     * reaches -1 and the reference's size_t conversion is undefined (its
     * histogram access lands before the array).  Such a v counts in bin 0,
     * and a quotient that rounds up to nbins at the top edge in the last bin. */
    size_t operator()(double v) const {
        const double q = (v - min) / bs;
        if (!(q > 0.)) return 0;
        const size_t b = (size_t)q;
        return b < nbins ? b : nbins - 1;
    }
    double bin_min(size_t b) const { return min + bs * (double)b; }
    bool too_low(double v) const { return (int)(v + 0.5) < (int)(min + 0.5); }
    bool too_high(double v) const { return (int)(v + 0.5) >= (int)(max + 0.5); }
    bool in_range(double v) const {
        return (int)(v + 0.5) >= (int)(min + 0.5) && (int)(v + 0.5) < (int)(max + 0.5);
    }
};

/* ---------------- glazier EQUAL_WINDOWS (glazier.hpp:69-115) ----------- */
struct WindowGeom { double min, max, bs; size_t nbins; };
std::vector<WindowGeom> construct_windows(double gmin, double gmax, double gbs, size_t nw, double ow) {
    std::vector<WindowGeom> w(nw);
    const double denom = 1 + (double)(nw - 1) * (1 - ow);
    const double wsize = (gmax - gmin) / denom;
    for (size_t k = 0; k < nw; ++k) {
        double wmin = gmin + (double)k * (1 - ow) * wsize;
        double wmax = wmin + wsize;
        if (k != 0)
            wmin = w[k - 1].min + gbs * (double)((size_t)((wmin - w[k - 1].min) / gbs));
        size_t nb = (size_t)std::ceil((wmax - wmin) / gbs);
        wmax = wmin + gbs * (double)nb;
        while (wmax > gmax) wmax -= gbs;
        w[k].min = wmin; w[k].max = wmax; w[k].bs = gbs;
        w[k].nbins = (size_t)((wmax - wmin) / gbs);
    }
    return w;
}

/* ---------------- model and walker ----------------------------------------- */
struct System;

struct Walker {
    int sys = 0, window = 0, replica = 0, local = 0;
    std::vector<double> s;                 /* Ising: N spins; AT: [s0,t0,s1,t1,...] */
    double E = 0, M = 0, Ms = 0, Mt = 0, Bax = 0;
    int which = 0;                         /* AT flavour: 0 sigma, 1 tau */
    HistIndex hi{};
    std::vector<double> lng;
    std::vector<uint64_t> cnt, visits;
    std::mt19937_64 gen;
    std::uniform_real_distribution<double> ureal{0., 1.};
    uint64_t exch_cursor = 0;
    uint64_t replay_draws = 0;             /* engine outputs consumed (replay) */
    std::vector<double> obs;               /* running means [bin][ob] (observables on) */
    bool obs_corr = false;                 /* observables 2: with G(qmin)           */
    std::vector<double> obs_counts;        /* counts_per_bin, as the reference keeps it */
    std::vector<uint32_t> dens;            /* AT sigma-tau density counts [bin][1024] (densities on) */
};

struct Window {
    HistIndex hi{};
    size_t nbins = 0;
    std::vector<uint8_t> exclude;
    double lnf = 1.0;
    int64_t iter = 1, sweeps = 0;
    bool oot = false, done = false;
};

struct System {
    rewl_system_desc d{};
    bool retired = false;                  /* rewl_params.retire: all windows done   */
    int L = 0, N = 0, ndof = 0;
    std::vector<double> field;             /* per-site field, empty = uniform h   */
    bool has_field = false;                /* AT: whether the sigma*tau term exists */
    std::vector<int> nbr;                  /* 4 per site in the reference order     */
    std::vector<Window> windows;
    int R = 1, W = 1;
    int first = 0;                         /* global index of local walker 0 */
    uint32_t k0 = 0, k1 = 0;
    double field_at(int i) const { return field.empty() ? d.h : field[i]; }
};

} // namespace

struct ora_ctx {
    std::vector<System> sys;
    std::vector<Walker> w;
    rewl_params p{};
    uint64_t sweep_index = 0;              /* sweeps done by every walker          */
    int direction_even = 1;                /* next exchange phase                  */
    uint64_t exchanges_accepted = 0, exchange_calls = 0, checks = 0, flat_events = 0;
};

namespace {

/* Neighbour order (x,y-1),(x-1,y),(x+1,y),(x,y+1) with site = y*L + x
 * (square_2d_grid_helpers.hpp:37-63, square_2d_nearest_neighbor_functor.cpp:90-110). */
void build_neighbors(System& S) {
    const int L = S.L;
    S.nbr.resize(4 * (size_t)S.N);
    for (int site = 0; site < S.N; ++site) {
        int x = site % L, y = site / L;
        S.nbr[4 * site + 0] = (y == 0 ? L - 1 : y - 1) * L + x;
        S.nbr[4 * site + 1] = y * L + (x == 0 ? L - 1 : x - 1);
        S.nbr[4 * site + 2] = y * L + (x == L - 1 ? 0 : x + 1);
        S.nbr[4 * site + 3] = (y == L - 1 ? 0 : y + 1) * L + x;
    }
}

/* Ising2d local energetics (ising2d_hamiltonian.hpp:162-231). */
inline double ising_local_energy(const System& S, const Walker& w, int i, double sv) {
    double f = S.field_at(i);
    const int* nb = &S.nbr[4 * (size_t)i];
    for (int k = 0; k < 4; ++k) f += S.d.J * w.s[nb[k]];
    return -1. * sv * f;
}

/* Ashkin_Teller2d local energetics (ashkin_teller2d_hamiltonian.hpp:169-191). */
inline double at_local_energy(const System& S, const Walker& w, int i, double sg, double ta) {
    double en = 0.;
    if (S.has_field) en += S.field_at(i) * (sg * ta);
    const int* nb = &S.nbr[4 * (size_t)i];
    for (int k = 0; k < 4; ++k) {
        const double sn = w.s[2 * (size_t)nb[k]], tn = w.s[2 * (size_t)nb[k] + 1];
        en += S.d.J * sg * sn;
        en += S.d.J * ta * tn;
        en += S.d.K * sg * ta * sn * tn;
    }
    return -en;
}

void recalculate_state(const System& S, Walker& w) {
    if (S.d.model == REWL_ISING2D) {
        double E = 0., M = 0.;
        for (int i = 0; i < S.N; ++i) {
            M += w.s[i];
            E += -0.5 * S.field_at(i) * w.s[i] + 0.5 * ising_local_energy(S, w, i, w.s[i]);
        }
        w.E = E; w.M = M;
    } else {
        double E = 0., sm = 0., tm = 0., bx = 0.;
        for (int i = 0; i < S.N; ++i) {
            const double sg = w.s[2 * (size_t)i], ta = w.s[2 * (size_t)i + 1];
            sm += sg; tm += ta; bx += sg * ta;
            if (S.has_field) E += -0.5 * S.field_at(i) * sg * ta;
            E += 0.5 * at_local_energy(S, w, i, sg, ta);
        }
        w.E = E; w.Ms = sm; w.Mt = tm; w.Bax = bx;
    }
}

struct Proposal { double E, M, Ms, Mt, Bax, sg, ta; };

/* change_state: the proposal for site i; for AT it also advances the flavour
 * after site N-1 (ashkin_teller2d_hamiltonian.hpp:223-250). */
inline Proposal change_state(const System& S, Walker& w, int i, bool toggle_at_end = true) {
    Proposal t{};
    if (S.d.model == REWL_ISING2D) {
        const double cur = w.s[i];
        t.sg = -cur;
        t.M = w.M + t.sg - cur;
        t.E = w.E + ising_local_energy(S, w, i, t.sg) - ising_local_energy(S, w, i, cur);
    } else {
        const double cs = w.s[2 * (size_t)i], ct = w.s[2 * (size_t)i + 1];
        double ts = cs, tt = ct;
        ts += -2. * ts * (double)(w.which == 0);
        tt += -2. * tt * (double)(w.which == 1);
        t.sg = ts; t.ta = tt;
        t.Ms = w.Ms + ts - cs;
        t.Mt = w.Mt + tt - ct;
        t.Bax = w.Bax + (ts * tt) - (cs * ct);
        t.E = w.E + at_local_energy(S, w, i, ts, tt) - at_local_energy(S, w, i, cs, ct);
        if (toggle_at_end && i == S.N - 1) w.which ^= 1;
    }
    return t;
}

inline void set_state(const System& S, Walker& w, int i, const Proposal& t) {
    w.E = t.E;
    if (S.d.model == REWL_ISING2D) { w.M = t.M; w.s[i] = t.sg; }
    else { w.Ms = t.Ms; w.Mt = t.Mt; w.Bax = t.Bax; w.s[2 * (size_t)i] = t.sg; w.s[2 * (size_t)i + 1] = t.ta; }
}

inline void switch_flavor(const System& S, Walker& w) { if (S.d.model == REWL_ASHKIN_TELLER2D) w.which ^= 1; }

/* ---------------- random draws -------------------------------------------- */
inline double replay_uniform(Walker& w) { ++w.replay_draws; return w.ureal(w.gen); }

inline double init_uniform(ora_ctx& c, const System& S, Walker& w) {
    if (c.p.rng_mode == REWL_RNG_REPLAY) return replay_uniform(w);
    P4 o = philox4x32_10(P4{{0u, 0u, (uint32_t)DOMAIN_INIT, (uint32_t)w.local}}, S.k0, S.k1);
    return canonical_from_u64((uint64_t)o.v[0] | ((uint64_t)o.v[1] << 32));
}

/* Ising2d_Obs / Ashkin_Teller2d_Obs::update_observable_average
 * (ising2d_observables.hpp:134-145): avg = (value + counts avg) / (counts + 1). */
inline void obs_average(Walker& w, size_t bin, int nobs, int ob, double value) {
    double& avg = w.obs[bin * (size_t)nobs + (size_t)ob];
    const double counts = w.obs_counts[bin];
    avg = (value + counts * avg) / (counts + 1);
}

/* update_observables at the walker's current bin: Ising
 * ising2d_hamiltonian.hpp:234-256 (|m|, m^2, m^4), Ashkin-Teller
 * ashkin_teller2d_hamiltonian.hpp:267-318 (sigma, tau, phi, Baxter: |x|, x^2,
 * x^4 with phi = sqrt(sigma^2 + tau^2)); then counts_per_bin += 1. */
/* Fourier_Correlator::compute_correlator (fourier_correlator.hpp:65-118) over
 * the walker's FP64 spins: field = spin of `type` (Ising: the spin; AT: sigma
 * or tau), or sigma * tau when mixed; kernel cos/sin(qmin idx), qmin = 2 pi/L
 * (fourier_correlator.hpp:14-42). */
double compute_correlator(const System& S, const Walker& w, int type, bool mixed) {
    const int L = S.L;
    const int nt = S.d.model == REWL_ISING2D ? 1 : 2;
    const double qmin = 2. * acos(-1.) / (double)L;
    double cr_x = 0., ci_x = 0., cr_y = 0., ci_y = 0.;
    for (int idx = 0; idx != L; ++idx) {
        double tcx = 0., tkr = 0., tki = 0.;
        for (int jdx = 0; jdx != L; ++jdx) {
            double fv = w.s[(size_t)(idx * L + jdx) * nt + type];
            fv *= ((double)(!mixed) + (double)mixed * w.s[(size_t)(idx * L + jdx) * nt + (nt - 1)]);
            tcx += fv;
            cr_y += fv * cos(qmin * (double)jdx);
            ci_y += fv * sin(qmin * (double)jdx);
            tkr += (double)(idx == jdx) * cos(qmin * (double)jdx);
            tki += (double)(idx == jdx) * sin(qmin * (double)jdx);
        }
        cr_x += tcx * tkr;
        ci_x += tcx * tki;
    }
    cr_x = cr_x * cr_x + ci_x * ci_x;
    cr_y = cr_y * cr_y + ci_y * ci_y;
    return 0.5 * (cr_x + cr_y);
}

/* update_qmin_correlator (ising2d_observables.hpp:152-159): running mean over
 * the n = counts_per_bin / 1000 earlier samples. */
inline void corr_average(Walker& w, size_t bin, int nobs, int ob, double value, double n) {
    double& avg = w.obs[bin * (size_t)nobs + (size_t)ob];
    avg = (value + n * avg) / (n + 1);
}

/* update_densities (ashkin_teller_densities.hpp:11-34 with the constants of
 * ashkin_teller_densities_parameters.cxx:15-30): one count in the cell
 * axis(tau) * axis_bins + axis(sigma) of the bin's 32 x 32 histogram. */
inline uint32_t density_axis_index(double observable) {
    const double sigma_tau_max = 1.05;
    const double sigma_tau_min = -sigma_tau_max;
    const double density_binwidth = (sigma_tau_max - sigma_tau_min) / (double)REWL_DENSITY_AXIS_BINS;
    return (uint32_t)((observable - sigma_tau_min) / density_binwidth);
}
inline void update_densities(Walker& w, size_t bin, double sigma, double tau) {
    const size_t idx = (size_t)density_axis_index(tau) * REWL_DENSITY_AXIS_BINS + density_axis_index(sigma);
    ++w.dens[bin * REWL_DENSITY_CELLS + idx];
}

/* update_observables at the walker's current bin: Ising
 * ising2d_hamiltonian.hpp:234-256, Ashkin-Teller
 * ashkin_teller2d_hamiltonian.hpp:267-318; with correlators (observables 2)
 * the reference's CORRELATION_LENGTHS layout, G(qmin) sampled when
 * counts_per_bin % 1000 == 0 (the AT tau correlator repeats the sigma one, as
 * at ashkin_teller2d_hamiltonian.hpp:298-299); then counts_per_bin += 1. */
inline void update_observables(const System& S, Walker& w) {
    if (w.obs.empty()) return;
    const size_t b = w.hi(w.E);
    const double divide_N = 1. / (double)S.N;
    const bool corr = w.obs_corr;
    const double counts = w.obs_counts[b];
    const bool sample = corr && ((size_t)counts % 1000 == 0);
    const double n = (double)((size_t)counts / 1000);
    if (S.d.model == REWL_ISING2D) {
        const int nobs = corr ? REWL_NUM_OBS_ISING_CORR : REWL_NUM_OBS_ISING;
        const double m = divide_N * std::fabs(w.M);
        obs_average(w, b, nobs, 0, m);
        obs_average(w, b, nobs, 1, m * m);
        obs_average(w, b, nobs, 2, m * m * m * m);
        if (sample) corr_average(w, b, nobs, 3, divide_N * divide_N * compute_correlator(S, w, 0, false), n);
    } else {
        const double sv = divide_N * std::fabs(w.Ms), tv = divide_N * std::fabs(w.Mt), bv = divide_N * std::fabs(w.Bax);
        const double phi2 = sv * sv + tv * tv;
        const int nobs = corr ? REWL_NUM_OBS_AT_CORR : REWL_NUM_OBS_AT;
        const int k = corr ? 4 : 3;   // stride of the (|x|, x^2, x^4[, G]) groups
        obs_average(w, b, nobs, 0, sv);
        obs_average(w, b, nobs, 1, sv * sv);
        obs_average(w, b, nobs, 2, sv * sv * sv * sv);
        obs_average(w, b, nobs, k + 0, tv);
        obs_average(w, b, nobs, k + 1, tv * tv);
        obs_average(w, b, nobs, k + 2, tv * tv * tv * tv);
        obs_average(w, b, nobs, 2 * k + 0, std::sqrt(phi2));
        obs_average(w, b, nobs, 2 * k + 1, phi2);
        obs_average(w, b, nobs, 2 * k + 2, phi2 * phi2);
        obs_average(w, b, nobs, 3 * k + 0, bv);
        obs_average(w, b, nobs, 3 * k + 1, bv * bv);
        obs_average(w, b, nobs, 3 * k + 2, bv * bv * bv * bv);
        if (!w.dens.empty()) update_densities(w, b, divide_N * w.Ms, divide_N * w.Mt);
        if (sample) {
            const double g_s = divide_N * divide_N * compute_correlator(S, w, 0, false);
            const double g_t = g_s;
            const double g_b = divide_N * divide_N * compute_correlator(S, w, 0, true);
            corr_average(w, b, nobs, 3, g_s, n);
            corr_average(w, b, nobs, 7, g_t, n);
            corr_average(w, b, nobs, 11, g_s + g_t, n);
            corr_average(w, b, nobs, 15, g_b, n);
        }
    }
    w.obs_counts[b] += 1.;
}

/* ---------------- Wang-Landau update and sweep (wang_landau.hpp:49-120) ---- */
struct SweepDraws {           /* Philox positional stream of one walker, one sweep */
    uint32_t k0, k1, walker;
    uint64_t base;            /* n * B */
    uint64_t cached_block = ~(uint64_t)0;
    P4 out{};
    double at(int p) {
        uint64_t blk = base + (uint64_t)(p >> 2);
        if (blk != cached_block) {
            out = philox4x32_10(P4{{(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)DOMAIN_SWEEP, walker}}, k0, k1);
            cached_block = blk;
        }
        return (double)out.v[p & 3] * 0x1p-32;
    }
};

inline void wl_update(const System& S, Walker& w, const Window& win, int i, bool replay, SweepDraws* sd, int p) {
    Proposal t = change_state(S, w, i);
    size_t cur = w.hi(w.E);
    if (w.hi.in_range(t.E)) {
        const size_t nb = w.hi(t.E);
        float d = w.lng[nb] - w.lng[cur];
        bool accept = d <= 0.;
        if (!accept) {
            const double u = replay ? replay_uniform(w) : sd->at(p);
            accept = u < static_cast<double>(std::exp(static_cast<double>(-d)));  /* exp(float) resolves to C exp(double) */
        }
        if (accept) { set_state(S, w, i, t); cur = nb; }
    }
    w.cnt[cur] += 1;
    w.lng[cur] += win.lnf;
    update_observables(S, w);
}

void sweep_walker(ora_ctx& c, System& S, Walker& w, uint64_t sweep_index) {
    const Window& win = S.windows[w.window];
    const bool replay = c.p.rng_mode == REWL_RNG_REPLAY;
    SweepDraws sd{S.k0, S.k1, (uint32_t)w.local, sweep_index * (uint64_t)((S.ndof + 3) / 4)};
    const int nflav = S.ndof / S.N;
    int p = 0;
    for (int f = 0; f < nflav; ++f)
        for (int i = 0; i < S.N; ++i, ++p) wl_update(S, w, win, i, replay, &sd, p);
}

/* ---------------- range adjustment (rewl_walker.hpp:71-115) --------------- */
int adjust_walker(ora_ctx& c, System& S, Walker& w) {
    bool in = w.hi.in_range(w.E);
    size_t site = (size_t)(init_uniform(c, S, w) * S.N);
    uint64_t guard = 0;
    const uint64_t cap = (uint64_t)S.ndof * 4096ull;
    while (!in) {
        if (++guard > cap) return fail(REWL_ERR_RANGE, "walker could not be brought into its window");
        ++site;
        if (site == (size_t)S.N) { site = 0; switch_flavor(S, w); }
        /* adjust_mode 1: only the wrap switches the flavour, so passes alternate */
        Proposal t = change_state(S, w, (int)site, c.p.adjust_mode == 0);
        if (w.hi.too_low(w.E) && t.E > w.E) set_state(S, w, (int)site, t);
        else if (w.hi.too_high(w.E) && t.E <= w.E) set_state(S, w, (int)site, t);
        in = w.hi.in_range(w.E);
    }
    return REWL_OK;
}

/* ---------------- flatness (rewl_histograms.hpp:115-152) ------------------ */
bool walker_flat(const ora_ctx& c, const Window& win, const Walker& w, const std::vector<uint8_t>& unseen) {
    if (c.p.schedule == REWL_SCHED_ONE_OVER_T && win.oot) return true;
    const size_t n = win.nbins;
    if (c.p.flatness_mode == REWL_FLAT_ALL_VISITED) {
        for (size_t b = 0; b < n; ++b)
            if (!win.exclude[b] && w.cnt[b] == 0) return false;
        return true;
    }
    uint64_t mn = UINT64_MAX, mx = 0, sum = 0; size_t ne = 0;
    for (size_t b = 0; b < n; ++b) {
        if (win.exclude[b]) continue;
        if (c.p.flatness_mode == 3 && unseen[b]) continue;   /* never visited by any replica */
        mn = std::min(mn, w.cnt[b]); mx = std::max(mx, w.cnt[b]); sum += w.cnt[b]; ++ne;
    }
    if (ne == 0) return true;
    if (c.p.flatness_mode == REWL_FLAT_RATIO) {
        float r = (mn != 0) ? (static_cast<float>(mx - mn) / static_cast<float>(mn)) : FLT_MAX;
        return r <= static_cast<float>(c.p.flatness);
    }
    if (mn == 0) return false;
    return (double)mn >= c.p.flatness * ((double)sum / (double)ne);
}

/* MPI_Allreduce(SUM) of one double per rank over a window communicator
 * (rewl_simulation.hpp:383, 416-429) as MPICH's recursive doubling computes
 * it, simulated rank by rank: with pof2 the largest power of two <= R and
 * rem = R - pof2, each odd rank below 2 rem adds the value of the even rank
 * before it and takes new rank r / 2, the ranks from 2 rem on take r - rem;
 * then, for mask = 1, 2, ..., pof2 / 2, every new rank adds its partner's
 * (new rank ^ mask) partial sum.  Every rank ends with the same value. */
double mpich_allreduce_sum(const std::vector<double>& v) {
    const int R = (int)v.size();
    int pof2 = 1;
    while (pof2 * 2 <= R) pof2 *= 2;
    const int rem = R - pof2;
    std::vector<double> part((size_t)pof2);
    for (int r = 0; r < R; ++r) {
        if (r < 2 * rem) {
            if (r % 2 == 1) part[(size_t)(r / 2)] = v[(size_t)(r - 1)] + v[(size_t)r];
        } else {
            part[(size_t)(r - rem)] = v[(size_t)r];
        }
    }
    for (int mask = 1; mask < pof2; mask <<= 1) {
        std::vector<double> next(part.size());
        for (int k = 0; k < pof2; ++k) next[(size_t)k] = part[(size_t)(k ^ mask)] + part[(size_t)k];
        part.swap(next);
    }
    return part[0];
}

/* count-weighted intra-window ln g average (rewl_simulation.hpp:361-454): the
 * weighted values summed over the replicas as the reference's MPI_Allreduce
 * sums them.  average_mode 1 weighs every replica 1/R (the alternative of
 * rewl_simulation.hpp:401). */
void average_window(ora_ctx& c, System& S, int wi, bool final_pass) {
    const int R = S.R;
    if (R == 1) return;
    Window& win = S.windows[wi];
    Walker* rep = &c.w[S.first + wi * R];
    std::vector<double> t((size_t)R);
    for (size_t b = 0; b < win.nbins; ++b) {
        if (c.p.average_mode == 1) {
            const double wgt = 1. / static_cast<double>(R);
            for (int r = 0; r < R; ++r) t[(size_t)r] = wgt * rep[r].lng[b];
        } else if (!final_pass) {
            uint64_t tot = 0;
            for (int r = 0; r < R; ++r) tot += rep[r].cnt[b];
            if (tot == 0) continue;
            for (int r = 0; r < R; ++r)
                t[(size_t)r] = (static_cast<double>(rep[r].cnt[b]) / static_cast<double>(tot)) * rep[r].lng[b];
        } else {
            std::vector<double> n((size_t)R);
            for (int r = 0; r < R; ++r) n[(size_t)r] = (double)(rep[r].visits[b] + rep[r].cnt[b]);
            const double tot = mpich_allreduce_sum(n);
            if (tot == 0.) continue;
            for (int r = 0; r < R; ++r) t[(size_t)r] = (n[(size_t)r] / tot) * rep[r].lng[b];
        }
        const double acc = mpich_allreduce_sum(t);
        for (int r = 0; r < R; ++r) rep[r].lng[b] = acc;
    }
}

/* reset_counts with REDUCE_LOGDOS (rewl_histograms.hpp:83-95) */
void reset_counts(const ora_ctx& c, Walker& w) {
    const double reducer = -w.lng[0] + 0.0;
    for (size_t b = 0; b < w.cnt.size(); ++b) {
        w.visits[b] += w.cnt[b];
        w.cnt[b] = 0;
        if (c.p.reduce_logdos) w.lng[b] += reducer;
    }
}

inline void update_histograms(System& S, Walker& w) {
    const size_t b = w.hi(w.E);
    w.cnt[b] += 1;
    w.lng[b] += S.windows[w.window].lnf;
}


void swap_states(Walker& a, Walker& b) {
    std::swap(a.E, b.E); std::swap(a.M, b.M); std::swap(a.Ms, b.Ms);
    std::swap(a.Mt, b.Mt); std::swap(a.Bax, b.Bax);
    std::swap(a.s, b.s);
}

/* one replica-exchange phase of one system (rewl_simulation.hpp:228-346) */
void exchange_system(ora_ctx& c, System& S, bool even) {
    const int R = S.R;
    for (int wl = even ? 0 : 1; wl + 1 < S.W; wl += 2) {
        Walker* lo = &c.w[S.first + wl * R];
        Walker* up = &c.w[S.first + (wl + 1) * R];
        std::vector<int> shuffled(R);
        for (int k = 0; k < R; ++k) shuffled[k] = R + k;
        if (c.p.rng_mode == REWL_RNG_REPLAY) {
            /* std::shuffle on the engine consumes raw 64-bit outputs */
            struct Counting {
                using result_type = uint64_t;
                Walker* w;
                static constexpr result_type min() { return std::mt19937_64::min(); }
                static constexpr result_type max() { return std::mt19937_64::max(); }
                result_type operator()() { ++w->replay_draws; return w->gen(); }
            } urbg{&lo[0]};
            std::shuffle(shuffled.begin(), shuffled.end(), urbg);
        } else {
            PhiloxExchangeUrbg urbg{S.k0, S.k1, (uint32_t)lo[0].local, &lo[0].exch_cursor};
            std::shuffle(shuffled.begin(), shuffled.end(), urbg);
        }
        for (int r = 0; r < R; ++r) {
            Walker& a = lo[r];
            Walker& b = up[shuffled[r] - R];
            const double Ea = a.E, Eb = b.E;
            double pa = 0., pb = 0.;
            if (a.hi.in_range(Eb)) pa = static_cast<double>(exp(std::min(0., a.lng[a.hi(Ea)] - a.lng[a.hi(Eb)])));
            if (b.hi.in_range(Ea)) pb = static_cast<double>(exp(std::min(0., b.lng[b.hi(Eb)] - b.lng[b.hi(Ea)])));
            double p = pa * pb;
            bool accept = false;
            if (p != 0.) {
                float u;
                if (c.p.rng_mode == REWL_RNG_REPLAY) u = (float)replay_uniform(a);
                else {
                    PhiloxExchangeUrbg g{S.k0, S.k1, (uint32_t)a.local, &a.exch_cursor};
                    u = (float)canonical_from_u64(g());
                }
                accept = u < p;
            }
            if (accept) { swap_states(a, b); ++c.exchanges_accepted; }
            update_histograms(S, a);
            update_observables(S, a);
            update_histograms(S, b);
            update_observables(S, b);
        }
    }
}


} // namespace

/* =========================== extern "C" API ================================ */
extern "C" {

const char* ora_last_error(void) { return g_err.c_str(); }

/* Simulated_Annealer::simulate_annealing (simulated_annealer.hpp:92-194) for
 * `chains` independent chains of every system: randomize_dofs
 * (ising2d_hamiltonian.hpp:136-150, ashkin_teller2d_hamiltonian.hpp:146-155)
 * from Philox (ctr {j / 4, 0, 5, chain}, word j mod 4, u < 0.5 gives +1),
 * then per temperature step k = 0..num_temperatures at
 * T = Ti (Tf / Ti)^(k / num) (calculate_temperature, :18-22) blocks of
 * block_size sweeps of sa_update (:66-86; Philox ctr {lo(a / 4), hi(a / 4),
 * 4, chain}, word a mod 4 for attempt a, u = word 2^-32) until the block
 * statistics settle (:129-158, simulated_annealing_utilities.hpp), or
 * max_blocks blocks.  min_energy_found starts at +inf (uninitialised there). */
inline bool sa_not_within_error(double value, double avg, double stdev) {
    return !(avg - stdev <= value && value <= avg + stdev);
}
/* The reference's unqualified abs on a double resolves to ::abs(int) with the
 * headers it includes: the difference is truncated to int; out of int's range
 * the x86 conversion gives INT_MIN. */
inline bool sa_not_within_tolerance(double new_v, double old_v, double tol) {
    const double d = new_v - old_v;
    const double a = std::fabs(d) < 2147483648.0 ? (double)std::abs((int)d) : -2147483648.0;
    return a / old_v > tol;
}
inline void sa_array_statistics(const std::vector<double>& e, double* average, double* stdev) {
    *average = 0.;
    *stdev = 0.;
    for (double v : e) *average += v;
    *average /= (double)e.size();
    for (double v : e) *stdev += (v - *average) * (v - *average);
    *stdev = std::sqrt(*stdev / (double)e.size());
}

int ora_anneal(const rewl_system_desc* systems, int32_t nsys, const rewl_anneal_params* p, double* min_energy,
               double* final_energy, uint64_t* sweeps, int8_t* final_spins) {
    if (!systems || nsys < 1 || !p || p->num_temperatures < 1 || p->block_size < 1 || p->max_blocks < 1 || p->chains < 1)
        return fail(REWL_ERR_INVALID, "bad argument");
    int8_t* so = final_spins;
    for (int si = 0; si < nsys; ++si) {
        System S;
        S.d = systems[si];
        S.L = S.d.L; S.N = S.L * S.L;
        S.ndof = S.d.model == REWL_ISING2D ? S.N : 2 * S.N;
        if (S.d.field) S.field.assign(S.d.field, S.d.field + S.N);
        S.has_field = S.d.field != nullptr || S.d.h != 0.0;
        build_neighbors(S);
        const uint32_t k0 = (uint32_t)S.d.seed, k1 = (uint32_t)(S.d.seed >> 32);
        const int flavors = S.ndof / S.N;
        for (int ch = 0; ch < p->chains; ++ch) {
            Walker w;
            w.s.assign((size_t)S.ndof, 1.0);
            for (int j = 0; j < S.ndof; ++j) {
                const P4 r = philox4x32_10(P4{{(uint32_t)j >> 2, 0u, 5u, (uint32_t)ch}}, k0, k1);
                const double u = (double)r.v[j & 3] * 0x1p-32;
                w.s[(size_t)j] = u < 0.5 ? 1. : -1.;
            }
            recalculate_state(S, w);
            double min_found = std::numeric_limits<double>::infinity();
            uint64_t a = 0, total = 0;
            std::vector<double> equil((size_t)p->block_size);
            for (int k = 0; k <= p->num_temperatures; ++k) {
                const double T = p->initial_temperature *
                                 std::pow(p->final_temperature / p->initial_temperature,
                                          (double)k / (double)p->num_temperatures);
                const double beta = 1. / T;
                uint64_t sweep = 0;
                double mean = DBL_MAX, sd = 1.;
                bool incomplete = true;
                int blocks = 0;
                while (incomplete) {
                    for (int b = 0; b < p->block_size; ++b) {
                        for (int f = 0; f < flavors; ++f)
                            for (int site = 0; site < S.N; ++site) {
                                const Proposal t = change_state(S, w, site);
                                double paccept = t.E < w.E ? 1. : -1.;
                                if (paccept == -1.) paccept = std::exp(beta * (w.E - t.E));
                                const P4 r = philox4x32_10(P4{{(uint32_t)(a >> 2), (uint32_t)(a >> 34), 4u, (uint32_t)ch}},
                                                           k0, k1);
                                const double u = (double)r.v[a & 3] * 0x1p-32;
                                ++a;
                                if (u < paccept) set_state(S, w, site, t);
                                min_found = w.E < min_found ? w.E : min_found;
                            }
                        equil[(size_t)b] = w.E;
                    }
                    double avg = 0., sdn = 0.;
                    sa_array_statistics(equil, &avg, &sdn);
                    incomplete = sa_not_within_error(avg, mean, sd);
                    incomplete = incomplete || sa_not_within_tolerance(sdn * sdn, sd * sd, p->stdev_tolerance);
                    incomplete = (sdn != p->frozen_criterion) && incomplete;
                    mean = avg;
                    sd = sdn;
                    sweep += (uint64_t)p->block_size;
                    if (++blocks >= p->max_blocks) incomplete = false;
                }
                total += sweep;
            }
            const size_t o = (size_t)si * (size_t)p->chains + (size_t)ch;
            min_energy[o] = min_found;
            final_energy[o] = w.E;
            sweeps[o] = total;
            if (so)
                for (int j = 0; j < S.ndof; ++j) *so++ = w.s[(size_t)j] > 0 ? 1 : -1;
        }
    }
    return REWL_OK;
}

int ora_create(const rewl_system_desc* systems, int32_t nsys, const rewl_params* params, ora_ctx** out) {
    if (!systems || nsys <= 0 || !params || !out) return fail(REWL_ERR_INVALID, "null argument");
    ora_ctx* c = new ora_ctx();
    c->p = *params;
    int first = 0;
    for (int si = 0; si < nsys; ++si) {
        System S;
        S.d = systems[si];
        S.L = S.d.L; S.N = S.L * S.L;
        if (S.L < 2 || S.L > 64 || (S.d.model != REWL_ISING2D && S.d.model != REWL_ASHKIN_TELLER2D)) {
            delete c; return fail(REWL_ERR_INVALID, "bad model or L");
        }
        S.ndof = S.d.model == REWL_ISING2D ? S.N : 2 * S.N;
        if (S.d.field) S.field.assign(S.d.field, S.d.field + S.N);
        S.has_field = S.d.field != nullptr || S.d.h != 0.0;
        build_neighbors(S);
        S.R = S.d.replicas_per_window; S.W = S.d.num_windows;
        S.k0 = (uint32_t)S.d.seed; S.k1 = (uint32_t)(S.d.seed >> 32);
        auto geo = construct_windows(S.d.e_min, S.d.e_max, S.d.bin_size, (size_t)S.W, S.d.overlap);
        for (int k = 0; k < S.W; ++k) {
            Window win;
            win.hi = HistIndex{geo[k].min, geo[k].max, geo[k].bs, geo[k].nbins};
            win.nbins = geo[k].nbins;
            win.exclude.assign(win.nbins, 0);
            if (S.d.exclude_bins) {
                const size_t off = (size_t)std::llround((geo[k].min - S.d.e_min) / S.d.bin_size);
                for (size_t b = 0; b < win.nbins; ++b) win.exclude[b] = S.d.exclude_bins[off + b];
            }
            win.lnf = params->lnf_initial;
            S.windows.push_back(win);
        }
        S.first = first;
        for (int k = 0; k < S.W; ++k)
            for (int r = 0; r < S.R; ++r) {
                Walker w;
                w.sys = si; w.window = k; w.replica = r; w.local = k * S.R + r;
                w.s.assign((size_t)S.ndof, 1.0);
                w.hi = S.windows[k].hi;
                w.lng.assign(S.windows[k].nbins, 0.0);
                w.cnt.assign(S.windows[k].nbins, 0);
                w.visits.assign(S.windows[k].nbins, 0);
                if (c->p.observables) {
                    const bool corr = c->p.observables == 2;
                    const int nobs = S.d.model == REWL_ISING2D ? (corr ? REWL_NUM_OBS_ISING_CORR : REWL_NUM_OBS_ISING)
                                                                : (corr ? REWL_NUM_OBS_AT_CORR : REWL_NUM_OBS_AT);
                    w.obs_corr = corr;
                    w.obs.assign(S.windows[k].nbins * (size_t)nobs, 0.0);
                    w.obs_counts.assign(S.windows[k].nbins, 0.0);
                    if (c->p.densities) {
                        if (S.d.model != REWL_ASHKIN_TELLER2D) {
                            delete c;
                            return fail(REWL_ERR_INVALID, "densities are an Ashkin-Teller observable");
                        }
                        w.dens.assign(S.windows[k].nbins * (size_t)REWL_DENSITY_CELLS, 0u);
                    }
                }
                w.gen.seed(S.d.walker_seeds ? S.d.walker_seeds[w.local] : (uint64_t)1);
                recalculate_state(S, w);
                c->w.push_back(std::move(w));
            }
        first += S.W * S.R;
        c->sys.push_back(std::move(S));
    }
    *out = c;
    return REWL_OK;
}

int ora_destroy(ora_ctx* c) { delete c; return REWL_OK; }

int ora_adjust_to_range(ora_ctx* c) {
    for (auto& w : c->w) {
        int rc = adjust_walker(*c, c->sys[w.sys], w);
        if (rc) return rc;
    }
    return REWL_OK;
}

int ora_sweeps(ora_ctx* c, int32_t n) {
    for (int k = 0; k < n; ++k) {
        for (auto& w : c->w)
            if (!c->sys[w.sys].retired) sweep_walker(*c, c->sys[w.sys], w, c->sweep_index);
        ++c->sweep_index;
    }
    for (auto& S : c->sys)
        if (!S.retired)
            for (auto& win : S.windows) win.sweeps += n;
    return REWL_OK;
}

int ora_exchange(ora_ctx* c) {
    const bool even = c->direction_even != 0;
    for (auto& S : c->sys)
        if (!S.retired) exchange_system(*c, S, even);
    c->direction_even = !c->direction_even;
    ++c->exchange_calls;
    return REWL_OK;
}

int ora_check(ora_ctx* c, int32_t* all_done) {
    bool all = true;
    for (auto& S : c->sys) {
        if (S.retired) continue;   /* its windows are all done */
        for (int k = 0; k < S.W; ++k) {
            Window& win = S.windows[k];
            std::vector<uint8_t> unseen(win.nbins, 1);
            for (int r = 0; r < S.R; ++r) {
                const Walker& w = c->w[S.first + k * S.R + r];
                for (size_t b = 0; b < win.nbins; ++b)
                    if (w.visits[b] + w.cnt[b] != 0) unseen[b] = 0;
            }
            bool flat = true;
            for (int r = 0; r < S.R; ++r) flat = flat && walker_flat(*c, win, c->w[S.first + k * S.R + r], unseen);
            if (flat) {
                average_window(*c, S, k, false);
                for (int r = 0; r < S.R; ++r) reset_counts(*c, c->w[S.first + k * S.R + r]);
                if (c->p.schedule == REWL_SCHED_ONE_OVER_T && !win.oot)
                    win.oot = (win.lnf <= 1. / static_cast<double>(win.sweeps * win.iter));
                const bool oot = c->p.schedule == REWL_SCHED_ONE_OVER_T && win.oot;
                /* (0.5 f)(!oot) + (1/t)(oot) of the reference, written so that a
                 * check with no sweeps since the last flat event cannot give 0*inf */
                win.lnf = oot ? (0.5 * win.lnf) * 0.0 + 1. / static_cast<double>(win.iter * win.sweeps) : 0.5 * win.lnf;
                win.done = !(win.lnf >= c->p.lnf_final);
                ++win.iter; win.sweeps = 0;
                ++c->flat_events;
            }
            all = all && win.done;
        }
    }
    ++c->checks;
    if (c->p.retire)
        for (auto& S : c->sys) {
            bool done = true;
            for (const auto& win : S.windows) done = done && win.done;
            S.retired = S.retired || done;
        }
    if (all_done) *all_done = all;
    return REWL_OK;
}

int ora_run(ora_ctx* c, int64_t max_checks, int32_t* all_done) {
    const int per_check = c->p.sweeps_per_check / c->p.sweeps_per_exchange;
    int32_t done = 0;
    for (int64_t k = 0; k < max_checks && !done; ++k) {
        for (int u = 0; u < per_check; ++u) {
            ora_sweeps(c, c->p.sweeps_per_exchange);
            if (c->p.exchange) ora_exchange(c);
        }
        ora_check(c, &done);
    }
    if (all_done) *all_done = done;
    return done ? REWL_OK : REWL_ERR_ITER_CAP;
}

int ora_finalize(ora_ctx* c) {
    for (auto& S : c->sys)
        for (int k = 0; k < S.W; ++k) average_window(*c, S, k, true);
    return REWL_OK;
}

/* The engine's multicanonical refinement (rewl_refine), restated: the final
 * replica average, ln f = 0 everywhere, histograms cleared, `sweeps` sweeps
 * with an exchange after every sweeps_per_exchange, then ln g(b) += ln H(b)
 * per window (H summed over the replicas; unvisited bins unchanged). */
int ora_refine(ora_ctx* c, int64_t sweeps, int32_t sweeps_per_exchange) {
    for (auto& S : c->sys) S.retired = false;
    ora_finalize(c);
    for (auto& S : c->sys)
        for (auto& win : S.windows) win.lnf = 0.0;
    for (auto& w : c->w) {
        std::fill(w.visits.begin(), w.visits.end(), 0ull);
        std::fill(w.cnt.begin(), w.cnt.end(), 0ull);
    }
    int64_t done = 0, since = 0;
    while (done < sweeps) {
        const int64_t k = std::min<int64_t>({1000, (int64_t)sweeps_per_exchange - since, sweeps - done});
        ora_sweeps(c, (int32_t)k);
        done += k;
        since += k;
        if (since == sweeps_per_exchange) {
            if (c->p.exchange) ora_exchange(c);
            since = 0;
        }
        for (auto& w : c->w)   /* counts into visits after every launch, as the engine does */
            for (size_t b = 0; b < w.cnt.size(); ++b) { w.visits[b] += w.cnt[b]; w.cnt[b] = 0; }
    }
    for (auto& S : c->sys)
        for (int k = 0; k < S.W; ++k) {
            const Window& win = S.windows[k];
            for (size_t b = 0; b < win.nbins; ++b) {
                uint64_t H = 0;
                for (int r = 0; r < S.R; ++r) {
                    const Walker& w = c->w[S.first + k * S.R + r];
                    H += w.visits[b] + w.cnt[b];
                }
                if (H == 0) continue;
                const double v = c->w[S.first + k * S.R].lng[b] + std::log((double)H);
                for (int r = 0; r < S.R; ++r) c->w[S.first + k * S.R + r].lng[b] = v;
            }
        }
    return REWL_OK;
}

int32_t ora_num_walkers(const ora_ctx* c) { return (int32_t)c->w.size(); }

int ora_walker_layout(const ora_ctx* c, int32_t walker, int32_t* sys, int32_t* window, int32_t* replica, int32_t* nbins) {
    if (walker < 0 || walker >= (int32_t)c->w.size()) return fail(REWL_ERR_INVALID, "walker");
    const Walker& w = c->w[walker];
    *sys = w.sys; *window = w.window; *replica = w.replica; *nbins = (int32_t)w.lng.size();
    return REWL_OK;
}

int ora_window_geometry(const ora_ctx* c, int32_t sys, int32_t window, double* emin, double* emax, double* bs, int32_t* nbins) {
    const Window& win = c->sys[sys].windows[window];
    *emin = win.hi.min; *emax = win.hi.max; *bs = win.hi.bs; *nbins = (int32_t)win.nbins;
    return REWL_OK;
}

int ora_window_state(const ora_ctx* c, int32_t sys, int32_t window, double* lnf, int64_t* iter, int64_t* sweeps, int32_t* done, int32_t* oot) {
    const Window& win = c->sys[sys].windows[window];
    *lnf = win.lnf; *iter = win.iter; *sweeps = win.sweeps; *done = win.done; *oot = win.oot;
    return REWL_OK;
}

int ora_get_logdos(const ora_ctx* c, int32_t walker, double* out) {
    const Walker& w = c->w[walker];
    std::memcpy(out, w.lng.data(), w.lng.size() * sizeof(double));
    return REWL_OK;
}
int ora_get_counts(const ora_ctx* c, int32_t walker, uint64_t* out) {
    const Walker& w = c->w[walker];
    std::memcpy(out, w.cnt.data(), w.cnt.size() * sizeof(uint64_t));
    return REWL_OK;
}
int ora_get_visits(const ora_ctx* c, int32_t walker, uint64_t* out) {
    const Walker& w = c->w[walker];
    for (size_t b = 0; b < w.visits.size(); ++b) out[b] = w.visits[b];
    return REWL_OK;
}
/* running means [bin][ob] and counts_per_bin of one walker (observables on) */
int ora_get_observables(const ora_ctx* c, int32_t walker, double* means, double* counts) {
    if (!c || walker < 0 || walker >= (int32_t)c->w.size() || !means || !counts) return fail(REWL_ERR_INVALID, "bad argument");
    const Walker& w = c->w[walker];
    if (w.obs.empty()) return fail(REWL_ERR_INVALID, "observables are off");
    std::copy(w.obs.begin(), w.obs.end(), means);
    std::copy(w.obs_counts.begin(), w.obs_counts.end(), counts);
    return REWL_OK;
}

int ora_get_densities(const ora_ctx* c, int32_t walker, uint32_t* out) {
    if (!c || walker < 0 || walker >= (int32_t)c->w.size() || !out) return fail(REWL_ERR_INVALID, "bad argument");
    const Walker& w = c->w[walker];
    if (w.dens.empty()) return fail(REWL_ERR_INVALID, "densities are off");
    std::copy(w.dens.begin(), w.dens.end(), out);
    return REWL_OK;
}

int ora_get_spins(const ora_ctx* c, int32_t walker, int8_t* out) {
    const Walker& w = c->w[walker];
    for (size_t i = 0; i < w.s.size(); ++i) out[i] = w.s[i] > 0 ? 1 : -1;
    return REWL_OK;
}
int ora_get_state(const ora_ctx* c, int32_t walker, double* E, double* M, int32_t* which, uint64_t* draws) {
    const Walker& w = c->w[walker];
    *E = w.E;
    const System& S = c->sys[w.sys];
    *M = S.d.model == REWL_ISING2D ? w.M : w.Bax;
    *which = w.which;
    *draws = c->p.rng_mode == REWL_RNG_REPLAY ? w.replay_draws : w.exch_cursor;
    return REWL_OK;
}
int ora_recalculated_energy(ora_ctx* c, int32_t walker, double* E) {
    Walker copy = c->w[walker];
    recalculate_state(c->sys[copy.sys], copy);
    *E = copy.E;
    return REWL_OK;
}
int ora_set_logdos(ora_ctx* c, int32_t walker, const double* lng) {
    Walker& w = c->w[walker];
    std::memcpy(w.lng.data(), lng, w.lng.size() * sizeof(double));
    return REWL_OK;
}
int ora_set_window_lnf(ora_ctx* c, int32_t sys, int32_t window, double lnf) {
    c->sys[sys].windows[window].lnf = lnf;
    return REWL_OK;
}
int ora_set_spins(ora_ctx* c, int32_t walker, const int8_t* spins) {
    Walker& w = c->w[walker];
    for (size_t i = 0; i < w.s.size(); ++i) w.s[i] = spins[i] > 0 ? 1.0 : -1.0;
    recalculate_state(c->sys[w.sys], w);
    return REWL_OK;
}
void ora_stats(const ora_ctx* c, uint64_t* out) {
    out[0] = c->sweep_index; out[1] = c->exchange_calls; out[2] = c->exchanges_accepted;
    out[3] = c->checks; out[4] = c->flat_events;
}

/* Neighbours of a site in the reference order (test hook). */
int ora_neighbors(int32_t L, int32_t site, int32_t* out4) {
    System S; S.L = L; S.N = L * L; build_neighbors(S);
    for (int k = 0; k < 4; ++k) out4[k] = S.nbr[4 * (size_t)site + k];
    return REWL_OK;
}

/* Energy of an arbitrary spin configuration (test hook). spins: ndof int8. */
int ora_energy_of(const rewl_system_desc* d, const int8_t* spins, double* E) {
    System S; S.d = *d; S.L = d->L; S.N = S.L * S.L;
    S.ndof = d->model == REWL_ISING2D ? S.N : 2 * S.N;
    if (d->field) S.field.assign(d->field, d->field + S.N);
    S.has_field = d->field != nullptr || d->h != 0.0;
    build_neighbors(S);
    Walker w; w.s.resize(S.ndof);
    for (int i = 0; i < S.ndof; ++i) w.s[i] = spins[i] > 0 ? 1.0 : -1.0;
    recalculate_state(S, w);
    *E = w.E;
    return REWL_OK;
}

void ora_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    P4 o = philox4x32_10(P4{{ctr[0], ctr[1], ctr[2], ctr[3]}}, key[0], key[1]);
    for (int k = 0; k < 4; ++k) out[k] = o.v[k];
}

/* libstdc++ std::shuffle of [base, base+n) driven by mt19937_64(seed); writes
 * the permutation and the number of engine calls it consumed (test hook for
 * the engine's own restatement of the algorithm). */
int ora_shuffle_mt(uint64_t seed, int32_t n, int32_t base, int32_t* out, uint64_t* draws) {
    std::vector<int> v(n);
    for (int k = 0; k < n; ++k) v[k] = base + k;
    struct Counting {
        using result_type = uint64_t;
        std::mt19937_64 g; uint64_t n = 0;
        static constexpr result_type min() { return std::mt19937_64::min(); }
        static constexpr result_type max() { return std::mt19937_64::max(); }
        result_type operator()() { ++n; return g(); }
    } urbg{std::mt19937_64(seed)};
    std::shuffle(v.begin(), v.end(), urbg);
    for (int k = 0; k < n; ++k) out[k] = v[k];
    *draws = urbg.n;
    return REWL_OK;
}

void ora_mt19937_64(uint64_t seed, int64_t n, uint64_t* out) {
    std::mt19937_64 g(seed);
    for (int64_t k = 0; k < n; ++k) out[k] = g();
}

/* ------------------- results path ------------------------------------- */
/* 4th-order central / one-sided derivatives (numerical_derivatives.hpp:46-70) */
static double deriv(size_t i, const std::vector<double>& f, const std::vector<double>& a) {
    const size_t n = f.size();
    if (i == 0) return (f[1] - f[0]) / (a[1] - a[0]);
    if (i == n - 1) return (f[i] - f[i - 1]) / (a[i] - a[i - 1]);
    if (i == 1 || i == n - 2) return (f[i + 1] - f[i - 1]) / (2. * (a[i] - a[i - 1]));
    return (-f[i + 2] + 8. * f[i + 1] - 8. * f[i - 1] + f[i - 2]) / (12. * (a[i] - a[i - 1]));
}

/* The reference calls unqualified abs() on a double, which resolves to
 * int abs(int): the argument is truncated toward zero
 * (vector_matching.hpp:107, 116). */
static double trunc_abs(double v) { return (double)std::abs((int)v); }

/* concatenate_tables_multiple_overlap (concatenate_rewl_tables.hpp:167-305)
 * for ln g only, followed by normalize_logDoS (array_shift.hpp:15-43). */
int ora_stitch(const ora_ctx* c, int32_t sysi, double states_per_site, int32_t normalize,
               double* energies, double* logdos, int32_t capacity, int32_t* n_out) {
    const System& S = c->sys[sysi];
    std::vector<double> fe, fl;
    std::vector<std::vector<double>> et(S.W), lt(S.W);
    for (int k = 0; k < S.W; ++k) {
        const Walker& w = c->w[S.first + k * S.R];
        /* bin energies on the system's global grid: the same value for the same
         * bin in every window (min + bs b equals it whenever bs is dyadic, as in
         * the reference's integer bins; with other bin sizes the per-window sums
         * round differently and the reference's exact match would fail) */
        const long long g0 = std::llround((w.hi.min - S.d.e_min) / S.d.bin_size);
        for (size_t b = 0; b < w.lng.size(); ++b) {
            et[k].push_back(S.d.e_min + S.d.bin_size * (double)(g0 + (long long)b));
            lt[k].push_back(w.lng[b]);
        }
    }
    double shifter = 0.;
    for (size_t b = 0; b < et[0].size(); ++b) { fe.push_back(et[0][b]); fl.push_back(lt[0][b] + shifter); }
    for (int k = 1; k < S.W; ++k) {
        /* find_left_concatenation_indices (vector_matching.hpp:94-142) */
        const double right_e = et[k].front();
        long left_start = -1;
        for (long i = (long)fe.size() - 1; i != -1; --i) if (fe[i] == right_e) { left_start = i; break; }
        if (left_start < 0) return fail(REWL_ERR_INVALID, "windows do not overlap");
        const size_t nl = fl.size();
        double lb = deriv((size_t)left_start, fl, fe), rb = deriv(0, lt[k], et[k]);
        double bdiff = trunc_abs(lb - rb);
        size_t cl = (size_t)left_start;
        for (size_t idx = 1; idx != nl - (size_t)left_start; ++idx) {
            lb = deriv((size_t)left_start + idx, fl, fe);
            rb = deriv(idx, lt[k], et[k]);
            if ((trunc_abs(lb - rb) + 0.5) < (bdiff + 0.5)) { bdiff = trunc_abs(lb - rb); cl = (size_t)left_start + idx; }
        }
        size_t right_idx = 0;
        for (size_t idx = (size_t)left_start; idx != fe.size(); ++idx) {
            right_idx = idx - (size_t)left_start;
            if (idx < cl) {}
            else if (idx == cl) shifter = fl[idx] - lt[k][right_idx];
            else fl[idx] = lt[k][right_idx] + shifter;
        }
        for (size_t idx = right_idx + 1; idx < et[k].size(); ++idx) { fe.push_back(et[k][idx]); fl.push_back(lt[k][idx] + shifter); }
    }
    if (normalize) {
        double mx = 0.;
        for (double v : fl) mx = mx * (double)(mx > v) + v * (double)(mx <= v);
        double sum = 0.;
        for (double v : fl) sum += exp(-(mx - v));
        sum = log(sum);
        sum += mx;
        const double sh = static_cast<double>(S.N) * log(states_per_site) - sum - log(2.) * (double)(S.d.e_max == 0.);
        for (double& v : fl) v += sh;
    }
    *n_out = (int32_t)fl.size();
    for (size_t i = 0; i < fl.size() && (int32_t)i < capacity; ++i) { energies[i] = fe[i]; logdos[i] = fl[i]; }
    return REWL_OK;
}

/* Thermodynamics::calculate_thermodynamics, TRAPEZOIDAL_RULE energy part
 * (thermodynamics.hpp:77-236).  out[t*5 + {F,U,U2,S,C}]. */
int ora_thermodynamics(const double* E, const double* lng, int32_t nb, int32_t N, double tmin, double tmax, int32_t nT, double* out) {
    const double dT = (tmax - tmin) / static_cast<double>(nT);
    for (int32_t Ti = 1; Ti <= nT; ++Ti) {
        const int32_t t = nT - Ti;
        const double T = tmin + dT * static_cast<double>(t);
        auto expo = [&](int32_t b) { return lng[b] - (E[b] / T); };
        double mx = expo(nb - 1);
        for (int32_t b = 2; b <= nb; ++b) { double v = expo(nb - b); mx = (mx < v ? v : mx); }
        double part = 0.;
        for (int32_t b = 1; b <= nb; ++b) {
            const int32_t cb = nb - b;
            const double e = exp(expo(cb) - mx);
            part += (cb != 0 && cb != nb - 1) ? 2. * e : e;
        }
        part = 0.5 * part;
        double* o = out + 5 * (size_t)t;
        o[0] = -T * (mx + log(part)) / static_cast<double>(N);
        double U = 0., U2 = 0., Sx = 0.;
        for (int32_t b = 1; b <= nb; ++b) {
            const int32_t cb = nb - b;
            const double ev = E[cb];
            const double wgt = exp(expo(cb) - mx);
            if (cb == 0 || cb == nb - 1) { U += 0.5 * ev * wgt; U2 += 0.5 * ev * ev * wgt; Sx += 0.5 * lng[cb] * wgt; }
            else { U += ev * wgt; U2 += ev * ev * wgt; Sx += lng[cb] * wgt; }
        }
        U /= (N * part); U2 /= ((double)N * N * part); Sx /= (N * part);
        o[1] = U; o[2] = U2; o[3] = Sx;
        o[4] = N * (U2 - U * U) / static_cast<double>(T * T);
    }
    return REWL_OK;
}

/* The observable part of Thermodynamics::calculate_thermodynamics
 * (thermodynamics.hpp:156-236): per temperature, sum over bins (from the top
 * down) of O_b times the reduced Boltzmann weight, halved at the two end bins,
 * divided by the trapezoid partition function. */
int ora_thermo_observables(const double* E, const double* lng, const double* obs, int32_t nb, int32_t nobs,
                           double tmin, double tmax, int32_t nT, double* out) {
    const double dT = (tmax - tmin) / static_cast<double>(nT);
    for (int32_t Ti = 1; Ti <= nT; ++Ti) {
        const int32_t t = nT - Ti;
        const double T = tmin + dT * static_cast<double>(t);
        auto expo = [&](int32_t b) { return lng[b] - (E[b] / T); };
        double mx = expo(nb - 1);
        for (int32_t b = 2; b <= nb; ++b) { double v = expo(nb - b); mx = (mx < v ? v : mx); }
        double part = 0.;
        for (int32_t b = 1; b <= nb; ++b) {
            const int32_t cb = nb - b;
            const double e = exp(expo(cb) - mx);
            part += (cb != 0 && cb != nb - 1) ? 2. * e : e;
        }
        part = 0.5 * part;
        double* o = out + (size_t)nobs * t;
        for (int32_t ob = 0; ob < nobs; ++ob) o[ob] = 0.;
        for (int32_t b = 1; b <= nb; ++b) {
            const int32_t cb = nb - b;
            const double wgt = exp(expo(cb) - mx);
            for (int32_t ob = 0; ob < nobs; ++ob) {
                const double v = obs[(size_t)cb * nobs + ob];
                if (cb == 0 || cb == nb - 1) o[ob] += 0.5 * (v) * wgt;
                else o[ob] += (v) * wgt;
            }
        }
        for (int32_t ob = 0; ob < nobs; ++ob) o[ob] /= (part);
    }
    return REWL_OK;
}

/* ------------------- CPU baseline ------------------------------------- */
/* Sweeps every walker `sweeps` times with `threads` std::threads, walkers
 * partitioned round-robin; returns wall seconds and attempts. */
int ora_bench_sweeps(ora_ctx* c, int32_t sweeps, int32_t threads, double* seconds, uint64_t* attempts) {
    if (threads < 1) threads = 1;
    const uint64_t base = c->sweep_index;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([c, t, threads, sweeps, base]() {
            for (size_t k = (size_t)t; k < c->w.size(); k += (size_t)threads) {
                Walker& w = c->w[k];
                for (int s = 0; s < sweeps; ++s) sweep_walker(*c, c->sys[w.sys], w, base + (uint64_t)s);
            }
        });
    for (auto& th : pool) th.join();
    auto t1 = std::chrono::steady_clock::now();
    c->sweep_index += (uint64_t)sweeps;
    uint64_t att = 0;
    for (auto& w : c->w) att += (uint64_t)c->sys[w.sys].ndof * (uint64_t)sweeps;
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    *attempts = att;
    return REWL_OK;
}

} // extern "C"
