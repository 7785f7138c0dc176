// Device-side data layout and model energetics of the REWL engine.
//
// Layout in HBM (all lane-interleaved so that the 64 lanes of a wavefront touch
// 64 consecutive words):
//   * walkers are packed into groups of at most 64, one group per wavefront,
//     all walkers of a group belonging to one system (same L and model);
//   * lattice word (plane, row y) of the walker in lane j of group k lives at
//     lat[group.lat_base + (plane * L + y) * 64 + j]; bit x of the word is the
//     spin at site y*L + x, bit = 1 meaning spin -1 (so all-up is all zeros);
//   * histogram bin b of that walker lives at [group.hist_base + b * 64 + j] in
//     lng (FP64), cnt (u32) and visits (u64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.hpp"
#include "models.hpp"

namespace rewl {

constexpr int kLanes = 64;

struct DevSystem {
    int32_t model, L, N, ndof;
    int32_t intpath;        // integer energies (clean couplings, no field)
    int32_t twoJ, twoK;     // 2J, 2K on the integer path
    int32_t has_field;      // FP64 path: field term present
    double J, K, h;
    const double* field;    // per-site field (device), or nullptr for uniform h
    uint32_t k0, k1;        // Philox key
    int32_t first_walker, first_window, W, R;
};

struct DevWindow {
    double emin, emax, bs;
    int32_t nbins, sys, first_walker, R;
    int32_t imin;           // integer path: window minimum
    int32_t phi_lo, phi_hi; // phi(emin), phi(emax), phi(e) = (int)(e + 0.5)
    int32_t shift;          // integer path: log2(bin size)
    int32_t pad;
    int64_t excl_off;       // offset of this window's bins in the exclusion mask
};

struct DevGroup {
    int32_t sys, first_walker, nwalkers, nbins;
    int64_t hist_base, lat_base;
};

struct KArgs {
    const DevSystem* sys;
    const DevWindow* win;
    const DevGroup* grp;
    const int32_t* w_window;
    const int32_t* w_local;
    const int64_t* w_hist;
    const int64_t* w_lat;
    double* w_E;
    int32_t* w_flavor;
    uint64_t* w_cursor;
    double* win_lnf;
    int64_t* win_iter;
    int64_t* win_sweeps;
    int32_t* win_oot;
    int32_t* win_done;
    const int32_t* win_active;   // 0: the window's system has retired (rewl_params.retire)
    const uint8_t* exclude;
    uint64_t* lat;
    double* lng;
    uint32_t* cnt;
    uint64_t* visits;
    const uint64_t* ring;      // replay streams, ring_cap words per walker
    const uint64_t* ring_filled;
    uint64_t ring_cap;
    int32_t* err;
    unsigned long long* exch_accepted;
    const void* bound_tab;     // acceptance-bound table (accept_bound.hpp), kBoundTabN entries
    double* obs;               // observable sums [(hist index) * nobs + ob], or null
    uint32_t* obs_tot;         // with correlators: updates so far per (bin, walker), hist layout
    int32_t nobs;              // observables per bin slot (3 / 4 Ising, 12 / 16 with any AT system)
    int32_t obs_corr;          // 1: the layout holds correlator slots
    uint32_t* dens;            // AT sigma-tau densities [(hist index) * 1024 + cell], or null
    int32_t num_walkers, num_windows, num_groups;
};

// The reference's in-range test adds 0.5 and truncates toward zero
// (histogram_index.hpp:22-30); for integer energies this maps e -> e + (e<0).
__device__ __forceinline__ int phi_int(int e) { return e - (e >> 31); }

__device__ __forceinline__ int trunc_half(double e) { return (int)(e + 0.5); }

// Site (x, y) neighbours in the reference order: (x,y-1),(x-1,y),(x+1,y),(x,y+1)
// (square_2d_grid_helpers.hpp:37-63).
__device__ __forceinline__ void neighbors(int site, int L, int nb[4]) {
    const int x = site % L, y = site / L;
    nb[0] = (y == 0 ? L - 1 : y - 1) * L + x;
    nb[1] = y * L + (x == 0 ? L - 1 : x - 1);
    nb[2] = y * L + (x == L - 1 ? 0 : x + 1);
    nb[3] = (y == L - 1 ? 0 : y + 1) * L + x;
}

// w[p] of a per-plane register array, and its assignment, as selects over
// the planes (a runtime index into a register array would place the array in
// scratch memory).
template <int NP, class T>
__device__ __forceinline__ T plane_pick(const T (&w)[NP], int p) {
    T r = w[0];
#pragma unroll
    for (int q = 1; q < NP; ++q) r = p == q ? w[q] : r;
    return r;
}
template <int NP, class T, class V>
__device__ __forceinline__ void plane_set(T (&w)[NP], int p, V v) {
#pragma unroll
    for (int q = 0; q < NP; ++q) w[q] = p == q ? (T)v : w[q];
}

// Spin of plane p at site as +-1 double, read from the packed lattice.
__device__ __forceinline__ double spin_at(const uint64_t* lat, int64_t base, int L, int plane, int site) {
    const int x = site % L, y = site / L;
    const uint64_t w = lat[base + (int64_t)(plane * L + y) * kLanes];
    return ((w >> x) & 1ull) ? -1.0 : 1.0;
}

__device__ __forceinline__ double field_at(const DevSystem& S, int site) {
    return S.field ? S.field[site] : S.h;
}

// The spins of one site and of its four neighbours in the reference order
// (x,y-1),(x-1,y),(x+1,y),(x,y+1), every plane of model M, read from memory.
template <int M>
__device__ __forceinline__ void site_spins_mem(const DevSystem& S, const uint64_t* lat, int64_t base, int site,
                                               double own[Model<M>::kPlanes], double nb[Model<M>::kPlanes][4]) {
    int n[4];
    neighbors(site, S.L, n);
#pragma unroll
    for (int q = 0; q < Model<M>::kPlanes; ++q) {
        own[q] = spin_at(lat, base, S.L, q, site);
#pragma unroll
        for (int k = 0; k < 4; ++k) nb[q][k] = spin_at(lat, base, S.L, q, n[k]);
    }
}

// E' of the proposal at `site` with flavour fl, from the lattice in memory
// (change_state in the reference order).
template <int M>
__device__ __forceinline__ double proposal_energy_mem(const DevSystem& S, const uint64_t* lat, int64_t base, int site,
                                                      int fl, double E) {
    double own[Model<M>::kPlanes], nb[Model<M>::kPlanes][4];
    site_spins_mem<M>(S, lat, base, site, own, nb);
    return Model<M>::proposal_energy(S.J, S.K, S.has_field, field_at(S, site), E, own, nb, fl);
}

// Magnetisations of a walker's lattice (bit 1 = spin -1): m[0] = sum of
// plane 0 (Ising M, AT sigma), m[1] = sum of the last plane (AT tau), m[2] =
// sum of their product (AT Baxter).
template <int M>
__device__ __forceinline__ void lattice_magnetisations(const uint64_t* lat, int L, int m[3]) {
    constexpr int NP = Model<M>::kPlanes;
    int ps = 0, pt = 0, px = 0;
    for (int y = 0; y < L; ++y) {
        const uint64_t s = lat[(int64_t)y * kLanes];
        const uint64_t t = lat[(int64_t)((NP - 1) * L + y) * kLanes];
        ps += __popcll(s);
        pt += __popcll(t);
        px += __popcll(s ^ t);
    }
    const int N = L * L;
    m[0] = N - 2 * ps;
    m[1] = N - 2 * pt;
    m[2] = N - 2 * px;
}

// update_observables at bin b (ising2d_hamiltonian.hpp:234-256,
// ashkin_teller2d_hamiltonian.hpp:267-318): the observable values of the
// current state, added to the walker's per-bin sums (walker-private, so the
// adds need no ordering; the no-return atomic keeps them off the walker's
// dependency chain).
__device__ __forceinline__ void obs_add(double* p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// update_observables at a bin (the model's values, Model<M>::values), added
// to the bin's sums: value i goes to slot (i / 3) * k + i % 3, k = 3, or 4
// when the layout leaves a correlator slot after each group of three (the
// reference's CORRELATION_LENGTHS enum order).
template <int M>
__device__ __forceinline__ void update_observables(double* obs, int N, const int m[3], bool corr = false) {
    double v[Model<M>::kValues];
    Model<M>::values(N, m, v);
    const int k = corr ? 4 : 3;
#pragma unroll
    for (int i = 0; i < Model<M>::kValues; ++i) obs_add(obs + (i / 3) * k + i % 3, v[i]);
}

// update_densities (ashkin_teller_densities.hpp:17-34, called at
// ashkin_teller2d_hamiltonian.hpp:291-293): the cell of (sigma/N, tau/N) in the
// 32 x 32 grid over [-1.05, 1.05]^2, counted in the walker's histogram of
// this bin.  The axis index truncates (x - min) / binwidth with the reference's
// FP64 constants (min = -1.05, binwidth = 2.1 / 32, an exact quotient).
__device__ __forceinline__ int density_axis(double x) {
    constexpr double tmax = 1.05;
    constexpr double tmin = -tmax;
    constexpr double width = (tmax - tmin) / 32.;
    return (int)(uint32_t)((x - tmin) / width);
}
__device__ __forceinline__ void update_density(uint32_t* dens, int64_t hidx, int N, const int m[3]) {
    const double divide_N = 1. / (double)N;
    const int cell = density_axis(divide_N * (double)m[1]) * 32 + density_axis(divide_N * (double)m[0]);
    __hip_atomic_fetch_add(dens + hidx * 1024 + cell, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// G(qmin) of one +-1 field given as L bit rows (bit 1 = -1): the reference's
// Fourier_Correlator::compute_correlator (fourier_correlator.hpp:65-118),
// 0.5 (|sum_y R_y e^{i q y}|^2 + |sum_x C_x e^{i q x}|^2), from row sums R
// (popcounts) and column sums C (bit-sliced vertical counters), divided by
// N^2 as update_observables does.  cs holds cos(q k), then sin(q k).
// The four Fourier sums of one +-1 field: f = {sum_y R_y cos(q y),
// sum_y R_y sin(q y), sum_x C_x cos(q x), sum_x C_x sin(q x)}, rows summed in
// order (the reference's order, fourier_correlator.hpp:65-118).
template <typename RowFn, typename CsFn>
__device__ void fourier_parts(int L, RowFn row, CsFn cs, double f[4]) {
    uint64_t c[7] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};   // count of set bits per column, bit-sliced
    double rr = 0., ri = 0.;
    for (int y = 0; y < L; ++y) {
        const uint64_t w = row(y);
        const double Ry = (double)(L - 2 * (int)__popcll(w));   // __popcll is unsigned
        rr += Ry * cs(y);
        ri += Ry * cs(L + y);
        uint64_t carry = w;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const uint64_t t = c[k] & carry;
            c[k] ^= carry;
            carry = t;
        }
    }
    double cr = 0., ci = 0.;
    for (int x = 0; x < L; ++x) {
        int n = 0;
#pragma unroll
        for (int k = 0; k < 7; ++k) n |= (int)((c[k] >> x) & 1ull) << k;
        const double Cx = (double)(L - 2 * n);
        cr += Cx * cs(x);
        ci += Cx * cs(L + x);
    }
    f[0] = rr; f[1] = ri; f[2] = cr; f[3] = ci;
}

// G(qmin) from the four Fourier sums, divided by N^2
__device__ __forceinline__ double fourier_g(int L, const double f[4]) {
    const double N = (double)L * (double)L;
    const double divide_N = 1. / N;
    return divide_N * divide_N * (0.5 * ((f[0] * f[0] + f[1] * f[1]) + (f[2] * f[2] + f[3] * f[3])));
}

template <typename RowFn>
__device__ double fourier_qmin(int L, RowFn row, const double* cs) {
    double f[4];
    fourier_parts(L, row, [&](int k) { return cs[k]; }, f);
    return fourier_g(L, f);
}

// update_qmin_correlator's samples (ising2d_hamiltonian.hpp:247-252,
// ashkin_teller2d_hamiltonian.hpp:296-316) added to the correlator slots of
// one bin where the model places them (Model<M>::add_correlators).  Row y of
// the lattice is taken from `cur` (the sweep's register copy), the others
// from HBM; y < 0 reads every row from HBM.
template <int M>
__device__ __noinline__ void add_correlators(double* obs, const uint64_t* lat, int L, int y, const uint64_t* cur,
                                             const double* cs) {
    constexpr int NP = Model<M>::kPlanes;
    auto plane = [&](int q, int r) -> uint64_t { return r == y ? cur[q] : lat[(int64_t)(q * L + r) * kLanes]; };
    const double gs = fourier_qmin(L, [&](int r) { return plane(0, r); }, cs);
    double gb = 0.;
    if (Model<M>::kBaxterCorrelator)
        gb = fourier_qmin(L, [&](int r) { return plane(0, r) ^ plane(NP - 1, r); }, cs);
    Model<M>::add_correlators([&](int slot, double g) { obs_add(obs + slot, g); }, gs, gb);
}

}  // namespace rewl
