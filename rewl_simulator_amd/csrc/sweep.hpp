// Wang-Landau sweep kernel: one lane per walker.
//
// Restates Wang_Landau::wang_landau_update / wang_landau_sweep and
// REWL_Walker::wang_landau_walk (wang_landau.hpp:49-120, rewl_walker.hpp:141-158):
// sites are visited in order 0..N-1 per flavour pass; a proposal outside the
// window is rejected without a draw; inside it, d = (float)(lng[new]-lng[cur])
// is accepted when d <= 0, otherwise a uniform is drawn and compared with
// exp(-d) in FP64; the count and ln g of the (possibly new) current bin are
// then incremented.
//
// Each wavefront owns one group of <= 64 walkers.  The group's ln g column
// ([bin][lane], FP64) and its visit counts live in LDS for the whole launch;
// the lattice streams row by row from HBM/L2 (one 64-bit word per row and
// plane, one row prefetched ahead), so every spin-flip attempt touches only
// registers and LDS.
#pragma once
#include "device.hpp"

namespace rewl {

// LDS arrays are [bin][ls] with ls the LDS lane stride (64, or the group's
// walker count when observables with correlators need more room per bin).
template <bool PACKED>
__device__ __forceinline__ void count_inc(uint32_t* s_cnt, int bin, int lane, int ls) {
    if (PACKED) {
        atomicAdd(&s_cnt[(bin >> 1) * ls + lane], 1u << ((bin & 1) << 4));
    } else {
        atomicAdd(&s_cnt[bin * ls + lane], 1u);
    }
}

template <bool PACKED>
__device__ __forceinline__ void flush_counts(uint32_t* s_cnt, uint32_t* g_cnt, int64_t hb, int nbins, int lane, int ls) {
    if (PACKED) {
        for (int b = 0; b < nbins; ++b) {
            const uint32_t v = (s_cnt[(b >> 1) * ls + lane] >> ((b & 1) << 4)) & 0xFFFFu;
            g_cnt[hb + (int64_t)b * kLanes] += v;
        }
        for (int b = 0; b < (nbins + 1) / 2; ++b) s_cnt[b * ls + lane] = 0u;
    } else {
        for (int b = 0; b < nbins; ++b) g_cnt[hb + (int64_t)b * kLanes] = s_cnt[b * ls + lane];
    }
}

// Uniform draw for one attempt.  Replay: next word of the walker's host-made
// mt19937_64 stream, mapped as uniform_real_distribution<double> does.
// Philox: word (p & 3) of the block drawn for positions p..p+3.
template <bool REPLAY>
__device__ __forceinline__ double draw_uniform(const uint64_t* ring, uint64_t cap, uint64_t filled, uint64_t& cursor,
                                               int32_t* err, const u32x4& rnd, int p) {
    if (REPLAY) {
        double u = 2.0;  // rejects if the stream ran dry (reported through err)
        if (cursor < filled) u = canonical_from_u64(ring[cursor & (cap - 1)]);
        else atomicOr(err, 1);
        ++cursor;
        return u;
    } else {
        const int k = p & 3;
        const uint32_t w = k == 0 ? rnd.x : (k == 1 ? rnd.y : (k == 2 ? rnd.z : rnd.w));
        return (double)w * 0x1p-32;
    }
}

// Acceptance core shared by every model on the integer path.  Returns true if
// the proposal was accepted; updates cb/lc/E and the LDS histogram column.
template <bool REPLAY, bool PACKED>
__device__ __forceinline__ bool wl_accept_int(int En, int& E, int& cb, double& lc, double lnf, int wmin, int plo,
                                              int pspan, int sh, double* s_lng, uint32_t* s_cnt, int lane, int ls,
                                              const uint64_t* ring, uint64_t cap, uint64_t filled, uint64_t& cursor,
                                              int32_t* err, const u32x4& rnd, int p) {
    const bool in = (unsigned)(phi_int(En) - plo) < (unsigned)pspan;
    // (size_t)((E - min) / bs) truncates toward zero: E = min - 1, which the
    // +0.5 range test lets through when min == 0, lands in bin 0
    const int nb = in ? max((En - wmin) >> sh, 0) : cb;
    const double ln = s_lng[nb * ls + lane];
    const float d = (float)(ln - lc);
    bool acc = in && (d <= 0.f);
    if (in && !(d <= 0.f)) {
        const double u = draw_uniform<REPLAY>(ring, cap, filled, cursor, err, rnd, p);
        acc = u < exp(-(double)d);
    }
    if (acc) { E = En; cb = nb; lc = ln; }
    lc += lnf;
    s_lng[cb * ls + lane] = lc;
    count_inc<PACKED>(s_cnt, cb, lane, ls);
    return acc;
}

// FP64 variant (field models): bins and range follow histogram_index.hpp:8-31,
// with the bin clamped to the window: the +0.5 range test admits energies down
// to -1.5 when the window minimum lies in (-0.5, 0.5), where a quotient <= -1
// makes the reference's size_t conversion undefined; such an energy counts in
// bin 0 (and a quotient rounding up to nbins at the top edge, in the last bin).
template <bool REPLAY, bool PACKED>
__device__ __forceinline__ bool wl_accept_fp(double En, double& E, int& cb, double& lc, double lnf, double emin,
                                             double bs, int plo, int phi, int nbmax, double* s_lng, uint32_t* s_cnt, int lane,
                                             int ls, const uint64_t* ring, uint64_t cap, uint64_t filled, uint64_t& cursor,
                                             int32_t* err, const u32x4& rnd, int p) {
    const int pe = trunc_half(En);
    const bool in = pe >= plo && pe < phi;
    const int nb = in ? min(max((int)((En - emin) / bs), 0), nbmax) : cb;
    const double ln = s_lng[nb * ls + lane];
    const float d = (float)(ln - lc);
    bool acc = in && (d <= 0.f);
    if (in && !(d <= 0.f)) {
        const double u = draw_uniform<REPLAY>(ring, cap, filled, cursor, err, rnd, p);
        acc = u < exp(-(double)d);
    }
    if (acc) { E = En; cb = nb; lc = ln; }
    lc += lnf;
    s_lng[cb * ls + lane] = lc;
    count_inc<PACKED>(s_cnt, cb, lane, ls);
    return acc;
}

__device__ __forceinline__ uint32_t bit(uint64_t w, int x) { return (uint32_t)(w >> x) & 1u; }
__device__ __forceinline__ double pm(uint32_t b) { return b ? -1.0 : 1.0; }


// MODEL: 0 Ising2d, 1 Ashkin-Teller2d.  INTP: integer energies.  OBS: per-bin
// observable sums (the walker tracks its magnetisations); OBS = 2 also samples
// the Fourier correlators at every 1000th update of a bin, from the lattice as
// it stands (rows above the current one already written back, the current one
// in registers), with each bin's update count in LDS (s_tot).
template <int MODEL, bool INTP, bool REPLAY, bool PACKED, int OBS>
__device__ void sweep_walker(const KArgs& a, const DevGroup& G, const DevSystem& S, int lane, int nsweeps,
                             uint64_t sweep_base, int flush_every, double* s_lng, uint32_t* s_cnt, uint32_t* s_tot,
                             const double* s_cs) {
    const int g = G.first_walker + lane;
    const int wi = a.w_window[g];
    const DevWindow Wn = a.win[wi];
    const double lnf = a.win_lnf[wi];
    const int64_t hb = G.hist_base + lane;
    const int nbins = G.nbins;
    const int ls = OBS == 2 ? G.nwalkers : kLanes;

    for (int b = 0; b < nbins; ++b) s_lng[b * ls + lane] = a.lng[hb + (int64_t)b * kLanes];
    if (PACKED) {
        for (int b = 0; b < (nbins + 1) / 2; ++b) s_cnt[b * ls + lane] = 0u;
    } else {
        for (int b = 0; b < nbins; ++b) s_cnt[b * ls + lane] = a.cnt[hb + (int64_t)b * kLanes];
    }
    if (OBS == 2)
        for (int b = 0; b < nbins; ++b) s_tot[b * ls + lane] = a.obs_tot[hb + (int64_t)b * kLanes];

    const int L = S.L;
    uint64_t* lat = a.lat + G.lat_base + lane;
    const uint32_t walker = (uint32_t)a.w_local[g];
    uint64_t cursor = REPLAY ? a.w_cursor[g] : 0ull;
    const uint64_t filled = REPLAY ? a.ring_filled[g] : 0ull;
    const uint64_t* ring = REPLAY ? a.ring + (int64_t)g * (int64_t)a.ring_cap : nullptr;
    const uint64_t cap = a.ring_cap;
    const uint64_t B = (uint64_t)((S.ndof + 3) / 4);
    int fl = a.w_flavor[g];

    int Ei = 0;
    double Ed = a.w_E[g];
    int cb;
    if (INTP) {
        Ei = (int)Ed;
        cb = max((Ei - Wn.imin) >> Wn.shift, 0);
    } else {
        cb = min(max((int)((Ed - Wn.emin) / Wn.bs), 0), Wn.nbins - 1);
    }
    double lc = s_lng[cb * ls + lane];
    const int pspan = Wn.phi_hi - Wn.phi_lo;
    int mag[3] = {0, 0, 0};
    double* obs = OBS ? a.obs + hb * a.nobs : nullptr;
    const int64_t obs_bin = (int64_t)kLanes * a.nobs;
    if (OBS) lattice_magnetisations<MODEL>(lat, S.L, mag);
    // The moments of consecutive updates in one bin are summed in registers
    // and added to the bin's sums when the walker changes bin (and at the end
    // of the launch): one atomic per moment per run instead of per update.
    constexpr int NV = Model<MODEL>::kValues;
    constexpr int GK = OBS == 2 ? 4 : 3;   // slot stride of the (|x|, x^2, x^4[, G]) groups
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.;
    int accb = -1;
    auto flush_acc = [&]() {
        if (accb < 0) return;
        double* p = obs + accb * obs_bin;
#pragma unroll
        for (int i = 0; i < NV; ++i) obs_add(p + (i / 3) * GK + i % 3, acc[i]);
    };
    // the observables of one update at bin b: the magnetisation moments, and
    // with OBS = 2 the correlators when the bin's count before it is 0 mod 1000
    auto observe = [&](int b, int y, const uint64_t* rows) {
        if (b != accb) {
            flush_acc();
            accb = b;
#pragma unroll
            for (int i = 0; i < NV; ++i) acc[i] = 0.;
        }
        double v[NV];
        Model<MODEL>::values(S.N, mag, v);
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[i] += v[i];
        if (Model<MODEL>::kDensities && a.dens) update_density(a.dens, hb + (int64_t)b * kLanes, S.N, mag);
        if (OBS == 2) {
            const uint32_t t = s_tot[b * ls + lane];
            s_tot[b * ls + lane] = t + 1u;
            if (t % 1000u == 0u) add_correlators<MODEL>(obs + b * obs_bin, lat, L, y, rows, s_cs);
        }
    };
    constexpr int NP = Model<MODEL>::kPlanes;
    constexpr int NPASS = Model<MODEL>::kPasses;

    for (int s = 0; s < nsweeps; ++s) {
        const uint64_t blk0 = (sweep_base + (uint64_t)s) * B;
        int p = 0;
        u32x4 rnd{0u, 0u, 0u, 0u};
        for (int pass = 0; pass < NPASS; ++pass) {
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
                for (int x = 0; x < L; ++x, ++p) {
                    if (!REPLAY && (p & 3) == 0) {
                        const uint64_t bi = blk0 + (uint64_t)(p >> 2);
                        rnd = philox4x32_10(u32x4{(uint32_t)bi, (uint32_t)(bi >> 32), DOMAIN_SWEEP, walker}, S.k0, S.k1);
                    }
                    const int xl = (x == 0) ? L - 1 : x - 1;
                    const int xr = (x == L - 1) ? 0 : x + 1;
                    // the bits (1 = -1) of the site and of its neighbours in the
                    // reference order, every plane
                    uint32_t ob[NP], nbb[NP][4];
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        ob[q] = bit(cu[q], x);
                        nbb[q][0] = bit(up[q], x);
                        nbb[q][1] = bit(cu[q], xl);
                        nbb[q][2] = bit(cu[q], xr);
                        nbb[q][3] = bit(dn[q], x);
                    }
                    const int P = Model<MODEL>::flip_plane(fl);
                    bool acc;
                    if (INTP) {
                        const int En = Ei + Model<MODEL>::int_energy_change(S.twoJ, S.twoK, ob, nbb, fl);
                        acc = wl_accept_int<REPLAY, PACKED>(En, Ei, cb, lc, lnf, Wn.imin, Wn.phi_lo, pspan, Wn.shift,
                                                           s_lng, s_cnt, lane, ls, ring, cap, filled, cursor, a.err, rnd, p);
                    } else {
                        double own[NP], nb[NP][4];
#pragma unroll
                        for (int q = 0; q < NP; ++q) {
                            own[q] = pm(ob[q]);
#pragma unroll
                            for (int k = 0; k < 4; ++k) nb[q][k] = pm(nbb[q][k]);
                        }
                        const double En =
                            Model<MODEL>::proposal_energy(S.J, S.K, S.has_field, field_at(S, y * L + x), Ed, own, nb, fl);
                        acc = wl_accept_fp<REPLAY, PACKED>(En, Ed, cb, lc, lnf, Wn.emin, Wn.bs, Wn.phi_lo, Wn.phi_hi,
                                                          Wn.nbins - 1, s_lng, s_cnt, lane, ls, ring, cap, filled, cursor,
                                                          a.err, rnd, p);
                    }
                    if (acc) plane_set(cu, P, plane_pick(cu, P) ^ (1ull << x));
                    if (OBS) {
                        Model<MODEL>::magnetise(mag, acc, (int)plane_pick(ob, P), (int)(ob[0] ^ ob[NP - 1]), fl);
                        observe(cb, y, cu);
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
            fl = Model<MODEL>::next_flavour(fl);   // flavour switches after site N-1
        }
        if (PACKED && ((s + 1) % flush_every) == 0) flush_counts<true>(s_cnt, a.cnt, hb, nbins, lane, ls);
    }

    if (PACKED) {
        if (nsweeps % flush_every) flush_counts<true>(s_cnt, a.cnt, hb, nbins, lane, ls);
    } else {
        flush_counts<false>(s_cnt, a.cnt, hb, nbins, lane, ls);
    }
    if (OBS) flush_acc();
    for (int b = 0; b < nbins; ++b) a.lng[hb + (int64_t)b * kLanes] = s_lng[b * ls + lane];
    if (OBS == 2)
        for (int b = 0; b < nbins; ++b) a.obs_tot[hb + (int64_t)b * kLanes] = s_tot[b * ls + lane];
    a.w_E[g] = INTP ? (double)Ei : Ed;
    a.w_flavor[g] = fl;
    if (REPLAY) a.w_cursor[g] = cursor;
}

// LDS: ln g [nbins][ls] FP64, counts [nbins][ls] u32 (or packed pairs), and
// with OBS = 2 update counts [nbins][ls] u32 and the cos/sin(qmin k) table.
template <int MODEL, bool INTP, bool REPLAY, bool PACKED, int OBS>
__global__ void __launch_bounds__(kLanes) k_sweep(KArgs a, const int32_t* groups, int ngroups, int nsweeps,
                                                  uint64_t sweep_base, int flush_every) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const DevGroup G = a.grp[groups[gi]];
        const DevSystem S = a.sys[G.sys];
        const int ls = OBS == 2 ? G.nwalkers : kLanes;
        double* s_lng = reinterpret_cast<double*>(smem);
        uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_lng + (size_t)G.nbins * ls);
        uint32_t* s_tot = s_cnt + (size_t)G.nbins * ls;
        double* s_cs = reinterpret_cast<double*>(smem + ((((size_t)G.nbins * ls * 16) + 15) & ~(size_t)15));
        if (OBS == 2) {
            // Transform_Kernel (fourier_correlator.hpp:14-42): cos and sin of qmin k
            const double qmin = 2. * acos(-1.) / (double)S.L;
            for (int k = lane; k < S.L; k += blockDim.x) {
                s_cs[k] = cos(qmin * (double)k);
                s_cs[S.L + k] = sin(qmin * (double)k);
            }
            __syncthreads();
        }
        if (lane < G.nwalkers)
            sweep_walker<MODEL, INTP, REPLAY, PACKED, OBS>(a, G, S, lane, nsweeps, sweep_base, flush_every, s_lng, s_cnt,
                                                           s_tot, s_cs);
        if (OBS == 2) __syncthreads();   // the table is rebuilt for the next group
    }
}

}  // namespace rewl
