// Wang-Landau sweep, Philox mode, FP64 energies (random-field Ising and
// random-strain Ashkin-Teller): the throughput kernel for field models.
//
// Same Markov chain, draws and FP64 arithmetic as k_sweep (and so as the
// reference's wang_landau_update, wang_landau.hpp:49-93, with the local
// energies of ising2d_hamiltonian.hpp:162-188 and
// ashkin_teller2d_hamiltonian.hpp:169-191 summed in the reference's order).
// What changes is what the walker wavefront waits for:
//
//  * the per-site field is staged in LDS once per launch, so its read is an
//    LDS access issued ahead of its use instead of a global load on the
//    dependency chain of every attempt;
//  * acceptance is the single comparison ln g[new] - ln g[cur] < B(k), with B
//    the FP64 bound of accept_bound.hpp for the attempt's Philox word k
//    (equivalent to d = (float)(ln g[new] - ln g[cur]) <= 0 or
//    u < exp(-(double)d) for a correctly rounded exp); producer wavefronts
//    compute B for every attempt of the next row into a double-buffered LDS
//    ring while the walker sweeps the current one (one barrier per row).
//
// LDS per group: ln g [nbins][stride] FP64, counts [nbins][stride] u32, the
// field [N] FP64, the ring: per walker two buffers of L FP64 bounds and one
// FP64 of padding (walker j's bound for position x of buffer b at
// ring[j (2L + 1) + b L + x]), so that a walker reads its bounds at constant
// offsets and the walkers' 8-byte reads fall on different banks.
#pragma once
#include "accept_bound.hpp"
#include "device.hpp"

namespace rewl {

__device__ __forceinline__ double pm_bit(uint64_t w, int x) { return ((w >> x) & 1ull) ? -1.0 : 1.0; }

// Bounds for every attempt position of one row of one sweep pass, for every
// walker of the group (producer lane j serves walker j); the producer
// wavefronts split the row's Philox blocks.
template <int MODEL>
__device__ void field_fill_row(const DevGroup& G, const DevSystem& S, int wave, int nwaves, int lane,
                               uint64_t sweep_base, int row, double* ring, const BoundEntry* tab) {
    constexpr int NPASS = Model<MODEL>::kPasses;
    const int L = S.L, stride = G.nwalkers;
    if (lane >= stride) return;
    const uint32_t walker = (uint32_t)(G.first_walker - S.first_walker + lane);
    const int per_sweep = NPASS * L;
    const int s = row / per_sweep, rem = row - s * per_sweep;
    const int pass = rem / L, y = rem - pass * L;
    const int p0 = pass * S.N + y * L;
    const int b0 = p0 >> 2, b1 = (p0 + L - 1) >> 2;
    const uint64_t B = (uint64_t)((S.ndof + 3) / 4);
    const uint64_t blk0 = (sweep_base + (uint64_t)s) * B;
    for (int blk = b0 + wave; blk <= b1; blk += nwaves) {
        const uint64_t bi = blk0 + (uint64_t)blk;
        const u32x4 r = philox4x32_10(u32x4{(uint32_t)bi, (uint32_t)(bi >> 32), DOMAIN_SWEEP, walker}, S.k0, S.k1);
        const int x0 = blk * 4 - p0;
        // all four bounds first (their table reads issue together), then the
        // positions of this row
        const double b4[4] = {accept_bound(r.x, tab), accept_bound(r.y, tab), accept_bound(r.z, tab),
                              accept_bound(r.w, tab)};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = x0 + j;
            if (x >= 0 && x < L) ring[lane * (2 * L + 1) + (row & 1) * L + x] = b4[j];
        }
    }
}

template <int MODEL>
__device__ void field_produce(const DevGroup& G, const DevSystem& S, int wave, int nwaves, int lane, int nsweeps,
                              uint64_t sweep_base, double* ring, const BoundEntry* tab) {
    constexpr int NPASS = Model<MODEL>::kPasses;
    const int rows = nsweeps * NPASS * S.L;
    field_fill_row<MODEL>(G, S, wave, nwaves, lane, sweep_base, 0, ring, tab);
    __syncthreads();   // pairs with the walker's barrier after loading the histograms
    for (int r = 0; r < rows; ++r) {
        if (r + 1 < rows) field_fill_row<MODEL>(G, S, wave, nwaves, lane, sweep_base, r + 1, ring, tab);
        __syncthreads();   // pairs with the walker's end-of-row barrier
    }
}

// OBS: the observables of update_observables (observables = 1) at every
// attempt, from magnetisations kept in registers; the moments of consecutive
// attempts in one bin are summed in registers and added to the bin's sums
// when the walker leaves it.
template <int MODEL, bool OBS = false>
__device__ void sweep_walker_field(const KArgs& a, const DevGroup& G, const DevSystem& S, int lane, int nsweeps,
                                   double* s_lng, uint32_t* s_cnt, const double* s_field, const double* ring) {
    const bool active = lane < G.nwalkers;
    const int g = G.first_walker + (active ? lane : 0);
    const int wi = a.w_window[g];
    const DevWindow Wn = a.win[wi];
    const double lnf = a.win_lnf[wi];
    const int64_t hb = G.hist_base + lane;
    const int nbins = G.nbins;
    const int stride = G.nwalkers;

    if (active)
        for (int b = 0; b < nbins; ++b) {
            s_lng[b * stride + lane] = a.lng[hb + (int64_t)b * kLanes];
            s_cnt[b * stride + lane] = a.cnt[hb + (int64_t)b * kLanes];
        }
    __syncthreads();   // histograms and field loaded, first row of bounds ready

    const int L = S.L;
    uint64_t* lat = a.lat + G.lat_base + lane;
    int fl = active ? a.w_flavor[g] : 0;
    double Ed = active ? a.w_E[g] : 0.0;
    int cb = min(max((int)((Ed - Wn.emin) / Wn.bs), 0), Wn.nbins - 1);
    double lc = active ? s_lng[cb * stride + lane] : 0.0;
    const int plo = Wn.phi_lo, phi = Wn.phi_hi, nbmax = Wn.nbins - 1;
    const double emin = Wn.emin, bs = Wn.bs;
    constexpr int NP = Model<MODEL>::kPlanes;
    constexpr int NPASS = Model<MODEL>::kPasses;
    int buf = 0;

    constexpr int NV = Model<MODEL>::kValues;
    int mag[3] = {0, 0, 0};
    double accv[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) accv[i] = 0.;
    int accb = -1;   // bin being summed (none yet)
    auto flush_obs = [&]() {
        if (accb < 0) return;
        double* p = a.obs + (hb + (int64_t)accb * kLanes) * a.nobs;
#pragma unroll
        for (int i = 0; i < NV; ++i) obs_add(p + i, accv[i]);   // observables = 1: slot i
    };
    if (OBS && active) lattice_magnetisations<MODEL>(lat, L, mag);

    for (int s = 0; s < nsweeps; ++s) {
        for (int pass = 0; pass < NPASS; ++pass) {
            uint64_t up[NP], cu[NP], dn[NP], r0[NP], nx[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                up[q] = active ? lat[(int64_t)(q * L + L - 1) * kLanes] : 0ull;
                cu[q] = active ? lat[(int64_t)(q * L) * kLanes] : 0ull;
                dn[q] = active ? lat[(int64_t)(q * L + (L > 1 ? 1 : 0)) * kLanes] : 0ull;
                r0[q] = 0ull;
            }
            for (int y = 0; y < L; ++y) {
              if (active) {
#pragma unroll
                for (int q = 0; q < NP; ++q) nx[q] = (y + 2 < L) ? lat[(int64_t)(q * L + y + 2) * kLanes] : 0ull;
                const double* brow = ring + lane * (2 * L + 1) + buf * L;
                const double* frow = s_field + y * L;
                for (int x = 0; x < L; ++x) {
                    const double B = brow[x];
                    const double hfield = frow[x];
                    const int xl = (x == 0) ? L - 1 : x - 1;
                    const int xr = (x == L - 1) ? 0 : x + 1;
                    // the spins of the site and of its neighbours in the
                    // reference order, every plane; E' by the model's
                    // change_state in the reference's order
                    double own[NP], nbs[NP][4];
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        own[q] = pm_bit(cu[q], x);
                        nbs[q][0] = pm_bit(up[q], x);
                        nbs[q][1] = pm_bit(cu[q], xl);
                        nbs[q][2] = pm_bit(cu[q], xr);
                        nbs[q][3] = pm_bit(dn[q], x);
                    }
                    const double En = Model<MODEL>::proposal_energy(S.J, S.K, S.has_field, hfield, Ed, own, nbs, fl);
                    // histogram_index.hpp:8-31, the bin clamped to the window
                    const int pe = trunc_half(En);
                    const bool in = pe >= plo && pe < phi;
                    const int nb = in ? min(max((int)((En - emin) / bs), 0), nbmax) : cb;
                    const double ln = s_lng[nb * stride + lane];
                    const bool acc = in && (ln - lc < B);
                    if (acc) { Ed = En; cb = nb; lc = ln; }
                    lc += lnf;
                    s_lng[cb * stride + lane] = lc;
                    atomicAdd(&s_cnt[cb * stride + lane], 1u);
                    if constexpr (OBS) {
                        // the flipped spin's bit before the flip (1 = -1), in its
                        // plane, and its Baxter bit
                        const int pb = (int)((plane_pick(cu, Model<MODEL>::flip_plane(fl)) >> x) & 1ull);
                        const int xb = (int)(((cu[0] ^ cu[NP - 1]) >> x) & 1ull);
                        Model<MODEL>::magnetise(mag, acc, pb, xb, fl);
                        if (cb != accb) {
                            flush_obs();
                            accb = cb;
#pragma unroll
                            for (int i = 0; i < NV; ++i) accv[i] = 0.;
                        }
                        double v[NV];
                        Model<MODEL>::values(S.N, mag, v);
#pragma unroll
                        for (int i = 0; i < NV; ++i) accv[i] += v[i];
                    }
                    if (acc) {
                        const int P = Model<MODEL>::flip_plane(fl);
                        plane_set(cu, P, plane_pick(cu, P) ^ (1ull << x));
                    }
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
              __syncthreads();   // this row's ring buffer consumed, the next one filled
              buf ^= 1;
            }
            fl = Model<MODEL>::next_flavour(fl);   // flavour switches after site N-1
        }
    }

    if (active) {
        if (OBS) flush_obs();
        for (int b = 0; b < nbins; ++b) {
            a.lng[hb + (int64_t)b * kLanes] = s_lng[b * stride + lane];
            a.cnt[hb + (int64_t)b * kLanes] = s_cnt[b * stride + lane];
        }
        a.w_E[g] = Ed;
        a.w_flavor[g] = fl;
    }
}

// A spin of a packed row as +-1.0 (bit 1 = -1): the high word of the double
// carries the bit as its sign, the low word is zero.  With x a constant (the
// row-unrolled walker) this is a bit extract and one shift-or.
__device__ __forceinline__ double spin_d(uint64_t w, int x) {
    const uint32_t half = x < 32 ? (uint32_t)w : (uint32_t)(w >> 32);
    // bit x moved to bit 31 and merged into the high word of 1.0: a shift
    // and one and-or
    const uint32_t hi = ((half << (31 - (x & 31))) & 0x80000000u) | 0x3FF00000u;
    return __hiloint2double((int)hi, 0);
}

// The row-unrolled field walker: the same chain, draws and FP64 arithmetic as
// sweep_walker_field, with L a template parameter so that every bit position
// is a constant and each spin of the current row is converted to +-1.0 once
// (as the right neighbour of the site before it) and carried through the
// row in registers, negated when it flips.
//
// The bin of a proposal is trunc(RN(RN(E' - min) / bs)) as the reference
// computes it (histogram_index.hpp:8-31); here the quotient is taken as
// y = RN(RN(E' - min) * RN(1 / bs)) and its truncation used directly unless y
// lies within y 2^-48 of an integer m >= 1, where the division is done (a
// branch no lane takes in practice).  Outside that margin RN(x / bs) lies on
// the same side of m as y: |y - x/bs| <= (x/bs) 2^-52 (1/bs rounded once,
// the product once) and |RN(x/bs) - x/bs| <= (x/bs) 2^-53, both below
// y 2^-49; y <= 0.5 truncates to 0 either way, and a negative quotient
// clamps to bin 0 either way.  So the bin equals the divided one, and the
// walker's LDS read no longer waits on a ten-deep FP64 division chain.
template <int MODEL, int L>
__device__ void sweep_walker_field_rows(const KArgs& a, const DevGroup& G, const DevSystem& S, int lane, int nsweeps,
                                        double* s_lng, uint32_t* s_cnt, const double* s_field, const double* ring) {
    using Mdl = Model<MODEL>;
    constexpr int NP = Mdl::kPlanes;
    constexpr int NPASS = Mdl::kPasses;
    const bool active = lane < G.nwalkers;
    const int g = G.first_walker + (active ? lane : 0);
    const int wi = a.w_window[g];
    const DevWindow Wn = a.win[wi];
    const double lnf = a.win_lnf[wi];
    const int64_t hb = G.hist_base + lane;
    const int nbins = G.nbins;
    const int stride = G.nwalkers;

    if (active)
        for (int b = 0; b < nbins; ++b) {
            s_lng[b * stride + lane] = a.lng[hb + (int64_t)b * kLanes];
            s_cnt[b * stride + lane] = a.cnt[hb + (int64_t)b * kLanes];
        }
    __syncthreads();   // histograms and field loaded, first row of bounds ready

    uint64_t* lat = a.lat + G.lat_base + lane;
    int fl = active ? a.w_flavor[g] : 0;
    double Ed = active ? a.w_E[g] : 0.0;
    int cb = min(max((int)((Ed - Wn.emin) / Wn.bs), 0), Wn.nbins - 1);
    double lc = active ? s_lng[__mul24(cb, stride) + lane] : 0.0;
    const int plo = Wn.phi_lo, phi = Wn.phi_hi, nbmax = Wn.nbins - 1;
    const double emin = Wn.emin, bs = Wn.bs;
    const double inv_bs = 1.0 / bs;
    const double J = S.J, K = S.K;
    int buf = 0;

    for (int s = 0; s < nsweeps; ++s) {
        for (int pass = 0; pass < NPASS; ++pass) {
            const int P = Mdl::flip_plane(fl);
            uint64_t up[NP], cu[NP], dn[NP], r0[NP], nx[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                up[q] = active ? lat[(int64_t)(q * L + L - 1) * kLanes] : 0ull;
                cu[q] = active ? lat[(int64_t)(q * L) * kLanes] : 0ull;
                dn[q] = active ? lat[(int64_t)(q * L + 1) * kLanes] : 0ull;
                r0[q] = 0ull;
            }
            for (int y = 0; y < L; ++y) {
                if (active) {
#pragma unroll
                    for (int q = 0; q < NP; ++q) nx[q] = (y + 2 < L) ? lat[(int64_t)(q * L + y + 2) * kLanes] : 0ull;
                    const double* brow = ring + lane * (2 * L + 1) + buf * L;
                    const double* frow = s_field + y * L;
                    // left neighbour, own spin of the current site (site 0's
                    // left is site L - 1, not yet visited)
                    double sl[NP], sc[NP];
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        sl[q] = spin_d(cu[q], L - 1);
                        sc[q] = spin_d(cu[q], 0);
                    }
#pragma unroll
                    for (int x = 0; x < L; ++x) {
                        // one site at a time: hoisting the row's loads and spin
                        // conversions ahead of the chain would hold L of each in
                        // registers
                        __builtin_amdgcn_sched_barrier(0);
                        const double B = brow[x];
                        const double hfield = frow[x];
                        // right neighbour: site x + 1, or site 0 of this row as it is now
                        double sr[NP], nbs[NP][4];
#pragma unroll
                        for (int q = 0; q < NP; ++q) {
                            sr[q] = spin_d(cu[q], x == L - 1 ? 0 : x + 1);
                            nbs[q][0] = spin_d(up[q], x);
                            nbs[q][1] = sl[q];
                            nbs[q][2] = sr[q];
                            nbs[q][3] = spin_d(dn[q], x);
                        }
                        // the system has a field (the walker is taken only then)
                        const double En = Mdl::proposal_energy(J, K, true, hfield, Ed, sc, nbs, fl);
                        const int pe = trunc_half(En);
                        const bool in = pe >= plo && pe < phi;
                        const double xe = En - emin;
                        const double yq = xe * inv_bs;
                        int nb = in ? min(max((int)yq, 0), nbmax) : cb;
                        double ln = s_lng[__mul24(nb, stride) + lane];
                        const double m = rint(yq);
                        if (in && yq > 0.5 && fabs(yq - m) <= yq * 0x1p-48) {
                            nb = min(max((int)(xe / bs), 0), nbmax);
                            ln = s_lng[__mul24(nb, stride) + lane];
                        }
                        const bool acc = in && (ln - lc < B);
                        if (acc) { Ed = En; cb = nb; lc = ln; }
                        lc += lnf;
                        s_lng[__mul24(cb, stride) + lane] = lc;
                        atomicAdd(&s_cnt[__mul24(cb, stride) + lane], 1u);
                        if (acc) {
                            plane_set(cu, P, plane_pick(cu, P) ^ (1ull << x));
                            plane_set(sc, P, -plane_pick(sc, P));
                        }
#pragma unroll
                        for (int q = 0; q < NP; ++q) {
                            sl[q] = sc[q];
                            sc[q] = sr[q];
                        }
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
                __syncthreads();   // this row's ring buffer consumed, the next one filled
                buf ^= 1;
            }
            fl = Mdl::next_flavour(fl);   // flavour switches after site N-1
        }
    }

    if (active) {
        for (int b = 0; b < nbins; ++b) {
            a.lng[hb + (int64_t)b * kLanes] = s_lng[b * stride + lane];
            a.cnt[hb + (int64_t)b * kLanes] = s_cnt[b * stride + lane];
        }
        a.w_E[g] = Ed;
        a.w_flavor[g] = fl;
    }
}

// One workgroup = one walker wavefront + NPROD producer wavefronts; groups are
// taken persistently in blockIdx order.  ROWS: systems with a field and L in
// {16, 24, 32, 48} take the row-unrolled walker (without observables; at
// L = 64 the unrolled Ashkin-Teller row would spill registers, so it keeps
// the loop, as do non-integral couplings without a field).
template <int MODEL, int NPROD, bool OBS = false, bool ROWS = false>
__global__ void __launch_bounds__(64 * (1 + NPROD)) k_sweep_field(KArgs a, const int32_t* groups, int ngroups,
                                                                  int nsweeps, uint64_t sweep_base) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const BoundEntry* tab = static_cast<const BoundEntry*>(a.bound_tab);
    // the walker wavefront issues whenever it is ready; a producer sharing its
    // SIMD fills the cycles it spends waiting
    if (tid < kLanes) __builtin_amdgcn_s_setprio(3);
    for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const DevGroup G = a.grp[groups[gi]];
        const DevSystem S = a.sys[G.sys];
        double* s_lng = reinterpret_cast<double*>(smem);
        uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_lng + (size_t)G.nbins * G.nwalkers);
        double* s_field = reinterpret_cast<double*>(smem + (((size_t)G.nbins * G.nwalkers * 12 + 15) & ~(size_t)15));
        double* ring = s_field + S.N;
        for (int i = tid; i < S.N; i += blockDim.x) s_field[i] = field_at(S, i);
        if (tid < kLanes) {
            if constexpr (ROWS && !OBS) {
                switch (S.has_field ? S.L : 0) {
                    case 16: sweep_walker_field_rows<MODEL, 16>(a, G, S, tid, nsweeps, s_lng, s_cnt, s_field, ring); break;
                    case 24: sweep_walker_field_rows<MODEL, 24>(a, G, S, tid, nsweeps, s_lng, s_cnt, s_field, ring); break;
                    case 32: sweep_walker_field_rows<MODEL, 32>(a, G, S, tid, nsweeps, s_lng, s_cnt, s_field, ring); break;
                    case 48: sweep_walker_field_rows<MODEL, 48>(a, G, S, tid, nsweeps, s_lng, s_cnt, s_field, ring); break;
                    default: sweep_walker_field<MODEL, OBS>(a, G, S, tid, nsweeps, s_lng, s_cnt, s_field, ring);
                }
            } else {
                sweep_walker_field<MODEL, OBS>(a, G, S, tid, nsweeps, s_lng, s_cnt, s_field, ring);
            }
        } else
            field_produce<MODEL>(G, S, tid / kLanes - 1, NPROD, tid & (kLanes - 1), nsweeps, sweep_base, ring, tab);
        __syncthreads();   // the ring, field and LDS histograms are reused by the next group
    }
}

}  // namespace rewl
