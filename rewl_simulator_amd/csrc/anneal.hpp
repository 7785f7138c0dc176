// Simulated annealing ground-state search (Simulated_Annealer,
// simulated_annealer.hpp:24-194): one lane per independent chain, 64 chains of
// one system per workgroup, the lattice bit-packed in HBM with the lanes of a
// group interleaved (the layout of the REWL walkers), the block energies of
// each chain in an HBM scratch array for the equilibration statistics.
#pragma once
#include <cfloat>

#include "aux_kernels.hpp"
#include "sweep.hpp"
#include "sweep_fast.hpp"

namespace rewl {

constexpr uint32_t DOMAIN_ANNEAL = 4;        // Metropolis draws
constexpr uint32_t DOMAIN_ANNEAL_INIT = 5;   // randomize_dofs
constexpr int kAnnealGroupWords = 2 * 64 * kLanes;   // two planes of 64 rows, 64 lanes

struct AnnealArgs {
    const DevSystem* sys;
    uint64_t* lat;             // [group][plane * L + y][lane]
    double* ebuf;              // [group][block][lane]: energies after each sweep of a block
    const double* beta;        // [num_temperatures + 1], 1 / T_k computed on the host
    int32_t num_temps, block_size, max_blocks, chains, groups_per_sys;
    double tol, frozen;
    double* min_energy;        // [system * chains + chain]
    double* final_energy;
    unsigned long long* sweeps;
};

// Double-double helpers for the exact Metropolis decision.
struct DD { double hi, lo; };
__device__ __forceinline__ DD dd_two_sum(double a, double b) {
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_mul(DD a, DD b) {
    const double p = a.hi * b.hi;
    const double e = fma(a.hi, b.hi, -p) + (a.hi * b.lo + a.lo * b.hi);
    const DD r = dd_two_sum(p, e);
    return r;
}
__device__ __forceinline__ DD dd_add_d(DD a, double b) {
    const DD s = dd_two_sum(a.hi, b);
    return dd_two_sum(s.hi, s.lo + a.lo);
}
// y / n for a double y and a small integer n, to double-double accuracy.
__device__ __forceinline__ DD dd_div_int(double y, int n) {
    const double q = y / (double)n;
    const double r = fma(-q, (double)n, y);
    return dd_two_sum(q, r / (double)n);
}

// exp(x) for x in [-23, 0] to ~1e-28 relative: exp(x / 1024) by its Taylor
// series (Horner, 14 terms), then squared ten times; returns the double
// nearest the double-double result.
__device__ __noinline__ double exp_rn_slow(double x) {
    const double y = x * 0x1p-10;   // exact
    DD s = {1.0, 0.0};
    for (int n = 14; n >= 1; --n) s = dd_add_d(dd_mul(dd_div_int(y, n), s), 1.0);
    for (int k = 0; k < 10; ++k) s = dd_mul(s, s);
    return s.hi + s.lo;
}

// u < RN(exp(x)) for x <= 0 and u = k 2^-32 in [0, 1): the comparison
// sa_update makes with a correctly rounded exp.  The device exp decides
// unless u lies within 2^-48 (relative) of it; then the double-double exp.
__device__ __forceinline__ bool exp_less(double u, double x) {
    if (x == 0.) return u < 1.;
    if (u == 0.) return x > -745.1332191019412;   // RN(exp(x)) > 0
    if (x < -23.) return false;                   // exp(x) < 2^-33 <= u
    const double p = exp(x);
    if (u < p * (1. - 0x1p-48)) return true;
    if (u > p * (1. + 0x1p-48)) return false;
    return u < exp_rn_slow(x);
}

// array_statistics (simulated_annealing_utilities.hpp:26-45): the mean, then
// the population standard deviation about it, in index order.
__device__ __forceinline__ void block_statistics(const double* e, int n, double& avg, double& sd) {
    double a = 0.;
    for (int i = 0; i < n; ++i) a += e[(int64_t)i * kLanes];
    a /= (double)n;
    double v = 0.;
    for (int i = 0; i < n; ++i) {
        const double d = e[(int64_t)i * kLanes] - a;
        v += d * d;
    }
    avg = a;
    sd = sqrt(v / (double)n);
}

// not_within_tolerance (simulated_annealing_utilities.hpp:21-25).  Its
// unqualified abs on a double resolves to ::abs(int) with the headers the
// reference includes, so the difference is truncated to int first; a value
// outside int's range converts to INT_MIN, as x86 does.
__device__ __forceinline__ bool sa_not_within_tolerance(double nv, double ov, double tol) {
    const double d = nv - ov;
    const double a = fabs(d) < 2147483648.0 ? (double)abs((int)d) : -2147483648.0;
    return a / ov > tol;
}

// One Metropolis attempt (sa_update, simulated_annealer.hpp:66-86): accept if
// the proposal lowers the energy, else if u < exp(beta (E - E')).
__device__ __forceinline__ bool metropolis(double E, double En, double beta, double u) {
    return En < E || exp_less(u, beta * (E - En));
}

// The lattice rows of the current site are held in registers as the REWL
// sweep holds them (the row above, the current row, the row below, row 0 as
// rewritten, the row after next), so an attempt reads no memory; local
// energies are summed in the reference's order (change_state).
template <int MODEL>
__global__ void __launch_bounds__(kLanes) k_anneal(AnnealArgs a) {
    extern __shared__ __align__(16) double s_field[];   // the system's per-site field, if any
    const int gid = blockIdx.x, lane = threadIdx.x;
    const int si = gid / a.groups_per_sys;
    const int chain = (gid % a.groups_per_sys) * kLanes + lane;
    const DevSystem S = a.sys[si];
    if (S.model != MODEL) return;   // the whole workgroup: the other model's launch anneals this system
    // every chain of the workgroup reads the same field: staged in LDS once,
    // so that an attempt does not wait on a global load
    if (S.field)
        for (int i = lane; i < S.N; i += kLanes) s_field[i] = S.field[i];
    __syncthreads();
    if (chain >= a.chains) return;   // no barrier follows
    uint64_t* lat = a.lat + (int64_t)gid * kAnnealGroupWords + lane;
    constexpr int NP = Model<MODEL>::kPlanes;
    const int L = S.L;

    // randomize_dofs: spin_array[j] = u < 0.5 ? +1 : -1, j in the reference
    // layout (AT: 2 site + flavour); bit 1 = spin -1
    {
        u32x4 r{0u, 0u, 0u, 0u};
        for (int y = 0; y < L; ++y) {
            uint64_t row[2] = {0ull, 0ull};
            for (int x = 0; x < L; ++x)
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    const uint32_t j = (uint32_t)(NP * (y * L + x) + q);   // < 8192
                    if ((j & 3u) == 0u)
                        r = philox4x32_10(u32x4{j >> 2, 0u, DOMAIN_ANNEAL_INIT, (uint32_t)chain}, S.k0, S.k1);
                    // u = w 2^-32 < 0.5 gives +1
                    if (word_of(r, (int)(j & 3u)) >= 0x80000000u) row[q] |= 1ull << x;
                }
#pragma unroll
            for (int q = 0; q < NP; ++q) lat[(int64_t)(q * L + y) * kLanes] = row[q];
        }
    }
    double E = recalc_energy<MODEL>(S, a.lat + (int64_t)gid * kAnnealGroupWords, lane);
    double emin = __builtin_inf();   // the reference leaves min_energy_found uninitialised
    unsigned long long total = 0ull;
    uint64_t ctr = 0ull;
    int fl = 0;
    u32x4 rnd{0u, 0u, 0u, 0u};
    double* eb = a.ebuf + (int64_t)gid * a.block_size * kLanes + lane;

    for (int t = 0; t <= a.num_temps; ++t) {
        const double beta = a.beta[t];
        double mean = DBL_MAX, sd = 1.;
        bool incomplete = true;
        int blocks = 0;
        while (incomplete) {
            for (int b = 0; b < a.block_size; ++b) {
                for (int pass = 0; pass < Model<MODEL>::kPasses; ++pass) {
                    uint64_t up[NP], cu[NP], dn[NP], r0[NP], nx[NP];
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        up[q] = lat[(int64_t)(q * L + L - 1) * kLanes];
                        cu[q] = lat[(int64_t)(q * L) * kLanes];
                        dn[q] = lat[(int64_t)(q * L + (L > 1 ? 1 : 0)) * kLanes];
                        r0[q] = 0ull;
                    }
                    for (int y = 0; y < L; ++y) {
#pragma unroll
                        for (int q = 0; q < NP; ++q) nx[q] = (y + 2 < L) ? lat[(int64_t)(q * L + y + 2) * kLanes] : 0ull;
                        for (int x = 0; x < L; ++x) {
                            if ((ctr & 3) == 0)
                                rnd = philox4x32_10(u32x4{(uint32_t)(ctr >> 2), (uint32_t)(ctr >> 34), DOMAIN_ANNEAL,
                                                          (uint32_t)chain}, S.k0, S.k1);
                            const double u = (double)word_of(rnd, (int)(ctr & 3)) * 0x1p-32;
                            ++ctr;
                            const int site = y * L + x;
                            const int xl = (x == 0) ? L - 1 : x - 1;
                            const int xr = (x == L - 1) ? 0 : x + 1;
                            // the spins as +-1 (every product of couplings and spins
                            // in the local energy is exact), summed in the reference
                            // order by the model's change_state
                            double own[NP], nb[NP][4];
#pragma unroll
                            for (int q = 0; q < NP; ++q) {
                                own[q] = pm(bit(cu[q], x));
                                nb[q][0] = pm(bit(up[q], x));
                                nb[q][1] = pm(bit(cu[q], xl));
                                nb[q][2] = pm(bit(cu[q], xr));
                                nb[q][3] = pm(bit(dn[q], x));
                            }
                            const double h = S.field ? s_field[site] : S.h;
                            const double En = Model<MODEL>::proposal_energy(S.J, S.K, S.has_field, h, E, own, nb, fl);
                            if (metropolis(E, En, beta, u)) {
                                const int P = Model<MODEL>::flip_plane(fl);
                                plane_set(cu, P, plane_pick(cu, P) ^ (1ull << x));
                                E = En;
                            }
                            emin = E < emin ? E : emin;
                        }
#pragma unroll
                        for (int q = 0; q < NP; ++q) {
                            lat[(int64_t)(q * L + y) * kLanes] = cu[q];
                            if (y == 0) r0[q] = cu[q];
                            up[q] = cu[q];
                            cu[q] = dn[q];
                            dn[q] = (y + 2 < L) ? nx[q] : r0[q];
                        }
                    }
                    fl = Model<MODEL>::next_flavour(fl);   // the flavour switches after site N - 1
                }
                eb[(int64_t)b * kLanes] = E;
            }
            double avg, sdn;
            block_statistics(eb, a.block_size, avg, sdn);
            incomplete = !(mean - sd <= avg && avg <= mean + sd);
            incomplete = incomplete || sa_not_within_tolerance(sdn * sdn, sd * sd, a.tol);
            incomplete = (sdn != a.frozen) && incomplete;
            mean = avg;
            sd = sdn;
            ++blocks;
            if (blocks >= a.max_blocks) incomplete = false;
        }
        total += (unsigned long long)blocks * (unsigned long long)a.block_size;
    }
    const int64_t o = (int64_t)si * a.chains + chain;
    a.min_energy[o] = emin;
    a.final_energy[o] = E;
    a.sweeps[o] = total;
}

}  // namespace rewl
