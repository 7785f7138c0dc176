// Wang-Landau sweep, Philox mode, integer energies: the throughput kernel.
//
// Same Markov chain, same draws and the same FP64 ln g arithmetic as k_sweep
// (and so as the reference's wang_landau_update, wang_landau.hpp:49-93); what
// changes is how each attempt is evaluated:
//
//  * neighbour disagreements: once per row, the static part (up, down and
//    right neighbours, none of which changes while the row is swept) is
//    bit-sliced into a 2-bit count per site; per attempt only the left
//    neighbour (the site just visited) is added.  The right neighbour of the
//    last site is site 0 of the same row and is corrected if it flipped.
//  * acceptance: u < exp(-d) is decided in the log domain.  T = -ln u is
//    computed from the Philox word in FP32 off the dependency chain; d < T - eps
//    accepts, d > T + eps rejects, and the rare attempts inside the band (and
//    u = 0) take the exact FP64 comparison u < exp(-(double)d).  eps = 2^-12
//    bounds the FP32 error of T with a wide margin, so every decision equals
//    the FP64 one.
//  * histograms: ln g (FP64) and u32 visit counts in LDS with the bin stride
//    equal to the number of walkers in the group, which the host picks so that
//    both fit.
#pragma once
#include "device.hpp"

namespace rewl {

constexpr float kLogBand = 0x1p-12f;
constexpr float kLn2 = 0.693147180559945309f;

// -ln(u) for u = k * 2^-32, NaN for k == 0 (forces the exact path).
__device__ __forceinline__ float neg_log_u32(uint32_t k) {
    const float kf = (float)k;
    const float t = (32.0f - __builtin_amdgcn_logf(kf)) * kLn2;   // v_log_f32 on kf >= 1
    return k == 0u ? __int_as_float(0x7fc00000) : t;
}

__device__ __forceinline__ uint32_t word_of(const u32x4& r, int k) {
    return k == 0 ? r.x : (k == 1 ? r.y : (k == 2 ? r.z : r.w));
}

__device__ __forceinline__ uint64_t rotr_row(uint64_t r, int L) {
    // bit x <- bit (x + 1) mod L
    return (r >> 1) | ((r & 1ull) << (L - 1));
}

template <int MODEL, bool SIMPLE, int NPROD>
__device__ __forceinline__ void wl_fast_attempt(int En, int& E, int& cbo, double& lc, double lnf, int wmin, int plo,
                                                int pspan, int sh, int stride8, int lane8, double* s_lng,
                                                uint32_t* s_cnt, float T, uint32_t k, uint64_t blk0, int p,
                                                uint32_t walker, uint32_t k0, uint32_t k1, bool& accepted) {
    bool in;
    int off;
    if (SIMPLE) {
        off = En - wmin;
        in = (unsigned)off < (unsigned)pspan;
    } else {
        in = (unsigned)(phi_int(En) - plo) < (unsigned)pspan;
        off = max(En - wmin, 0);
    }
    // byte offset of the proposed bin's ln g in LDS
    const int nbo = in ? (off >> sh) * stride8 + lane8 : cbo;
    const double ln = *(const double*)((const char*)s_lng + nbo);
    const float d = (float)(ln - lc);
    const float diff = d - T;
    bool acc = in && ((d <= 0.f) | (diff < -kLogBand));
    const bool exact = in && !(d <= 0.f) && !(diff < -kLogBand) && !(diff > kLogBand);
    if (exact) {
        if (NPROD) {
            const uint64_t bi = blk0 + (uint64_t)(p >> 2);
            k = word_of(philox4x32_10(u32x4{(uint32_t)bi, (uint32_t)(bi >> 32), DOMAIN_SWEEP, walker}, k0, k1), p & 3);
        }
        acc = (double)k * 0x1p-32 < exp(-(double)d);
    }
    if (acc) { E = En; cbo = nbo; lc = ln; }
    lc += lnf;
    *(double*)((char*)s_lng + cbo) = lc;
    atomicAdd((uint32_t*)((char*)s_cnt + (cbo >> 1)), 1u);
    accepted = acc;
}

// MODEL 0 Ising2d, 1 Ashkin-Teller2d; L <= 64; integer couplings; Philox draws.
// NPROD = 0: the walker wavefront draws its own Philox words.  NPROD > 0:
// NPROD producer wavefronts of the same workgroup write -ln(u) for every
// attempt of the next row into a double-buffered LDS ring (ring[2][L][stride])
// while the walkers sweep the current row; one barrier per row.
template <int MODEL, bool SIMPLE, int NPROD>
__device__ void sweep_walker_fast(const KArgs& a, const DevGroup& G, const DevSystem& S, int lane, int nsweeps,
                                  uint64_t sweep_base, double* s_lng, uint32_t* s_cnt, const float* ring) {
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
    if (NPROD) __syncthreads();   // histograms loaded, first row of thresholds ready

    const int L = S.L;
    uint64_t* lat = a.lat + G.lat_base + lane;
    const uint32_t walker = (uint32_t)a.w_local[g];
    const uint64_t B = (uint64_t)((S.ndof + 3) / 4);
    int fl = a.w_flavor[g];
    int E = (int)a.w_E[g];
    const int wmin = Wn.imin, sh = Wn.shift;
    const int pspan = SIMPLE ? (int)(Wn.emax - Wn.emin) : Wn.phi_hi - Wn.phi_lo;
    const int stride8 = stride * 8, lane8 = lane * 8;
    int cbo = (max(E - wmin, 0) >> sh) * stride8 + lane8;
    double lc = active ? *(const double*)((const char*)s_lng + cbo) : 0.0;
    int buf = 0;
    const uint64_t lmask = (L == 64) ? ~0ull : ((1ull << L) - 1ull);
    constexpr int NP = Model<MODEL>::kPlanes;
    constexpr int NPASS = Model<MODEL>::kPasses;
    constexpr bool KQ = Model<MODEL>::kProductCoupling;

    for (int s = 0; s < nsweeps; ++s) {
        const uint64_t blk0 = (sweep_base + (uint64_t)s) * B;
        u32x4 rnd{0u, 0u, 0u, 0u};
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
                // ---- per-row static disagreement counts --------------------------
                // P: the plane being flipped, Q: the other one (a one-plane
                // model has Q = P, unused)
                const int pf = Model<MODEL>::flip_plane(fl), qf = NP - 1 - pf;
                const uint64_t P = plane_pick(cu, pf), Pu = plane_pick(up, pf), Pd = plane_pick(dn, pf);
                const uint64_t Q = plane_pick(cu, qf), Qu = plane_pick(up, qf), Qd = plane_pick(dn, qf);
                const uint64_t pu = P ^ Pu, pd = P ^ Pd, pr = (P ^ rotr_row(P, L)) & lmask;
                const uint64_t c0 = pu ^ pd ^ pr, c1 = (pu & pd) | (pr & (pu ^ pd));
                uint64_t m0 = 0, m1 = 0, ql = 0;
                if constexpr (KQ) {
                    const uint64_t qu = Q ^ Qu, qd = Q ^ Qd, qr = (Q ^ rotr_row(Q, L)) & lmask;
                    const uint64_t xu = pu ^ qu, xd = pd ^ qd, xr = pr ^ qr;
                    m0 = xu ^ xd ^ xr;
                    m1 = (xu & xd) | (xr & (xu ^ xd));
                    // Q's left-neighbour disagreement: bit x = Q_x ^ Q_{x-1}
                    ql = (Q ^ ((Q << 1) | (Q >> (L - 1)))) & lmask;
                }
                uint64_t flips = 0ull;
                uint32_t sl = (uint32_t)(P >> (L - 1)) & 1u;   // left of site 0: site L-1, not yet visited
                const float* rrow = ring + (size_t)buf * L * stride + lane;
                const int pbase = pass * S.N + y * L;   // attempt position of site (0, y) in this sweep
                for (int x = 0; x < L; ++x) {
                    const int p = pbase + x;
                    uint32_t k = 0u;
                    float T;
                    if (NPROD) {
                        T = rrow[x * stride];
                    } else {
                        if ((p & 3) == 0) {
                            const uint64_t bi = blk0 + (uint64_t)(p >> 2);
                            rnd = philox4x32_10(u32x4{(uint32_t)bi, (uint32_t)(bi >> 32), DOMAIN_SWEEP, walker}, S.k0,
                                                S.k1);
                        }
                        k = word_of(rnd, p & 3);
                        T = neg_log_u32(k);
                    }
                    const uint32_t sP = (uint32_t)(P >> x) & 1u;
                    const uint32_t dl = sP ^ sl;
                    int cP = (int)((uint32_t)(c0 >> x) & 1u) + 2 * (int)((uint32_t)(c1 >> x) & 1u);
                    int En;
                    if constexpr (!KQ) {
                        if (x == L - 1) {
                            // right neighbour is site 0 of this row: undo the static guess if it flipped
                            const int f0 = (int)(flips & 1ull), dr = (int)((uint32_t)(pr >> x) & 1u);
                            cP += f0 * (1 - 2 * dr);
                        }
                        En = E + S.twoJ * (4 - 2 * (cP + (int)dl));
                    } else {
                        int cq = (int)((uint32_t)(m0 >> x) & 1u) + 2 * (int)((uint32_t)(m1 >> x) & 1u);
                        const uint32_t qlx = (uint32_t)(ql >> x) & 1u;
                        if (x == L - 1) {
                            const int f0 = (int)(flips & 1ull);
                            const int dr = (int)((uint32_t)(pr >> x) & 1u);
                            const int xr = dr ^ (int)(((Q ^ (Q >> 1) ^ (Q << (L - 1))) >> x) & 1u) ;
                            cP += f0 * (1 - 2 * dr);
                            cq += f0 * (1 - 2 * xr);
                        }
                        const int pf = cP + (int)dl;
                        const int qq = cq + (int)(dl ^ qlx);
                        En = E + S.twoJ * (4 - 2 * pf) + S.twoK * (4 - 2 * qq);
                    }
                    bool acc;
                    wl_fast_attempt<MODEL, SIMPLE, NPROD>(En, E, cbo, lc, lnf, wmin, Wn.phi_lo, pspan, sh, stride8,
                                                          lane8, s_lng, s_cnt, T, k, blk0, p, walker, S.k0, S.k1, acc);
                    flips |= (uint64_t)acc << x;
                    sl = sP ^ (uint32_t)acc;
                }
                plane_set(cu, pf, P ^ flips);
#pragma unroll
                for (int q = 0; q < NP; ++q) {
                    lat[(int64_t)(q * L + y) * kLanes] = cu[q];
                    if (y == 0) r0[q] = cu[q];
                    up[q] = cu[q];
                    cu[q] = dn[q];
                    dn[q] = (y + 2 < L) ? nx[q] : r0[q];
                }
              }
              if (NPROD) {
                  __syncthreads();   // this row's ring buffer consumed, the next one filled
                  buf ^= 1;
              }
            }
            fl = Model<MODEL>::next_flavour(fl);
        }
    }

    if (active) {
        for (int b = 0; b < nbins; ++b) {
            a.lng[hb + (int64_t)b * kLanes] = s_lng[b * stride + lane];
            a.cnt[hb + (int64_t)b * kLanes] = s_cnt[b * stride + lane];
        }
        a.w_E[g] = (double)E;
        a.w_flavor[g] = fl;
    }
}

// Producer side of the ring: -ln(u) for every attempt position of one row of
// one sweep pass, for every walker of the group, from the same Philox words
// the walkers would draw themselves.  Producer lane j serves walker j of the
// group; the NPROD producer wavefronts split the row's Philox blocks.
template <int MODEL>
__device__ void fill_row(const DevGroup& G, const DevSystem& S, int wave, int nwaves, int lane, uint64_t sweep_base,
                         int row, float* ring_buf) {
    constexpr int NPASS = Model<MODEL>::kPasses;
    const int L = S.L, stride = G.nwalkers;
    if (lane >= stride) return;
    // walkers of a group are consecutive local ids of one system
    const uint32_t walker = (uint32_t)(G.first_walker - S.first_walker + lane);
    const int per_sweep = NPASS * L;
    const int s = row / per_sweep, rem = row - s * per_sweep;
    const int pass = rem / L, y = rem - pass * L;
    const int p0 = pass * S.N + y * L;
    const int b0 = p0 >> 2, b1 = (p0 + L - 1) >> 2;
    const uint64_t B = (uint64_t)((S.ndof + 3) / 4);
    const uint64_t blk0 = (sweep_base + (uint64_t)s) * B;
    const bool aligned = ((p0 | L) & 3) == 0;
    for (int blk = b0 + wave; blk <= b1; blk += nwaves) {
        const uint64_t bi = blk0 + (uint64_t)blk;
        const u32x4 r = philox4x32_10(u32x4{(uint32_t)bi, (uint32_t)(bi >> 32), DOMAIN_SWEEP, walker}, S.k0, S.k1);
        const int x0 = blk * 4 - p0;
        if (aligned) {
            float* o = ring_buf + x0 * stride + lane;
            o[0] = neg_log_u32(r.x);
            o[stride] = neg_log_u32(r.y);
            o[2 * stride] = neg_log_u32(r.z);
            o[3 * stride] = neg_log_u32(r.w);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int x = x0 + j;
                if (x >= 0 && x < L) ring_buf[x * stride + lane] = neg_log_u32(word_of(r, j));
            }
        }
    }
}

template <int MODEL>
__device__ void produce_thresholds(const DevGroup& G, const DevSystem& S, int wave, int nwaves, int lane, int nsweeps,
                                   uint64_t sweep_base, float* ring) {
    constexpr int NPASS = Model<MODEL>::kPasses;
    const int L = S.L;
    const int rows = nsweeps * NPASS * L;
    const size_t bufsz = (size_t)L * G.nwalkers;
    fill_row<MODEL>(G, S, wave, nwaves, lane, sweep_base, 0, ring);
    __syncthreads();   // pairs with the walkers' barrier after loading the histograms
    for (int r = 0; r < rows; ++r) {
        if (r + 1 < rows) fill_row<MODEL>(G, S, wave, nwaves, lane, sweep_base, r + 1, ring + ((r + 1) & 1) * bufsz);
        __syncthreads();   // pairs with the walkers' end-of-row barrier
    }
}

// One workgroup = one walker wavefront (+ NPROD producer wavefronts); groups
// are taken persistently in blockIdx order.
template <int MODEL, bool SIMPLE, int NPROD>
__global__ void __launch_bounds__(256) k_sweep_fast(KArgs a, const int32_t* groups, int ngroups, int nsweeps,
                                                    uint64_t sweep_base) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const DevGroup G = a.grp[groups[gi]];
        const DevSystem S = a.sys[G.sys];
        double* s_lng = reinterpret_cast<double*>(smem);
        uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_lng + (size_t)G.nbins * G.nwalkers);
        float* ring = reinterpret_cast<float*>(s_cnt + (size_t)G.nbins * G.nwalkers);
        if (tid < kLanes)
            sweep_walker_fast<MODEL, SIMPLE, NPROD>(a, G, S, tid, nsweeps, sweep_base, s_lng, s_cnt, ring);
        else if (NPROD)
            produce_thresholds<MODEL>(G, S, tid / kLanes - 1, NPROD, tid & (kLanes - 1), nsweeps, sweep_base, ring);
        if (NPROD) __syncthreads();   // the ring and LDS histograms are reused by the next group
    }
}

}  // namespace rewl

namespace rewl {

// Largest |T(k) - (-ln(k 2^-32))| over k = base + i, i < n (k = 0 skipped),
// the error the acceptance band kLogBand has to cover.
__global__ void k_selfcheck_neglog(uint32_t base, uint32_t n, unsigned long long* out_bits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double err = 0.0;
    for (uint32_t j = i; j < n; j += gridDim.x * blockDim.x) {
        const uint32_t k = base + j;
        if (k == 0u) continue;
        const double t = (double)neg_log_u32(k);
        const double ref = -log((double)k * 0x1p-32);
        err = fmax(err, fabs(t - ref));
    }
    for (int o = 32; o > 0; o >>= 1) err = fmax(err, __shfl_xor(err, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, (unsigned long long)__double_as_longlong(err));
}

}  // namespace rewl
