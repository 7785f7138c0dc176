// Wang-Landau sweep, Philox mode, integer energies, lattice side L known at
// compile time: the row-unrolled throughput kernel.
//
// Same chain, same draws and the same FP64 ln g arithmetic as k_sweep and
// k_sweep_fast (the reference's wang_landau_update, wang_landau.hpp:49-93).
// With one walker per lane and only a few thousand walkers, a lane's time per
// attempt is set by the length of its instruction stream (a wavefront issues
// at most one instruction per SIMD slot), so this kernel is written to
// minimise instructions per attempt and has no branch inside a row:
//
//  * the site loop of a row is unrolled (L = 16, 24, 32, 48, 64), so every
//    bit position is a constant and neighbour bits come out of 32-bit row
//    halves with one bit-field extract each;
//  * per row, the static disagreement count cs (up, down, right neighbours)
//    of every site is packed as the signed 2-bit field cs - 2, and the left
//    neighbour as the static pdl = P ^ rotl(P) corrected by the previous
//    site's acceptance a, so that Delta E = -4J (cs - 2 + (pdl ^ a)) is two
//    extracts, an add and one 24-bit multiply-add;
//  * the window offset o = E - window minimum is the walker's state, so the
//    proposal's LDS slot comes straight from o + Delta E;
//  * guard bins around the window hold ln g = +inf, so a proposal outside the
//    window is rejected by the acceptance test itself, with no range test;
//  * producer wavefronts of the same workgroup turn every Philox word into
//    the bound B of accept_bound.hpp, so acceptance is the single FP64
//    comparison ln g[new] - ln g[cur] < B;
//  * ln g (FP64) and the u32 count of a (bin, walker) share one 16-byte LDS
//    slot, so the count increment is a ds_add at offset 8 of the ln g slot.
//
// With observables = 2 (the reference's default CORRELATION_LENGTHS build) the
// slot's last word holds the bin's update count before the launch minus its
// count, so that one 16-byte read of a candidate slot gives ln g and the
// bin's all-time update count; the Fourier sums of the correlator are kept in
// registers, moved by each row's flips at the end of the row, and a sample
// (every 1000th update of a bin) combines them with the row's flips so far.
#pragma once
#include "accept_bound.hpp"
#include "device.hpp"
#include "sweep_fast.hpp"

namespace rewl {

constexpr int kTunedSlot = 16;   // bytes per (bin, walker): f64 ln g, u32 count, pad

// bytes of one walker's bound ring: two buffers of L FP64 values + 16 pad
constexpr __host__ __device__ int tuned_ring_stride(int L) { return 16 * L + 16; }

__device__ __forceinline__ uint32_t bit_of(uint32_t w, int b) { return __builtin_amdgcn_ubfe(w, b, 1); }

// bit i of a 16-bit value to bit 2 i
__device__ __forceinline__ uint32_t spread16(uint32_t v) {
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

template <int L>
__device__ __forceinline__ uint64_t rotr1(uint64_t r) {   // bit x <- bit (x + 1) mod L
    return (r >> 1) | ((r & 1ull) << (L - 1));
}
template <int L>
__device__ __forceinline__ uint64_t rotl1(uint64_t r) {   // bit x <- bit (x - 1) mod L
    constexpr uint64_t m = L == 64 ? ~0ull : ((1ull << L) - 1ull);
    return ((r << 1) | (r >> (L - 1))) & m;
}

// Dynamic LDS of the tuned kernel.  Slots are addressed by their 32-bit LDS
// address (lds_base() + byte offset), so that an access is one ds_* with no
// address arithmetic left at the access.
extern __shared__ __align__(16) unsigned char tuned_lds[];

typedef __attribute__((address_space(3))) double lds_double;
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) unsigned char lds_uchar;
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) f64x2 lds_f64x2_t;

__device__ __forceinline__ uint32_t lds_base() {
    return (uint32_t)(uintptr_t)(lds_uchar*)(tuned_lds);
}
__device__ __forceinline__ double lds_f64(uint32_t addr) { return *(lds_double*)(uintptr_t)addr; }
__device__ __forceinline__ void lds_store_f64(uint32_t addr, double v) { *(lds_double*)(uintptr_t)addr = v; }
__device__ __forceinline__ f64x2 lds_f64x2(uint32_t addr) { return *(lds_f64x2_t*)(uintptr_t)addr; }
__device__ __forceinline__ void lds_store_f64x2(uint32_t addr, f64x2 v) { *(lds_f64x2_t*)(uintptr_t)addr = v; }
// ln g += ln f as an LDS FP64 atomic (no return): the same IEEE add as a
// load, add and store, in one instruction
__device__ __forceinline__ void lds_add_f64(uint32_t addr, double v) {
    __hip_atomic_fetch_add((lds_double*)(uintptr_t)addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4v lds_u32x4_t;
__device__ __forceinline__ u32x4v lds_u32x4(uint32_t addr) { return *(lds_u32x4_t*)(uintptr_t)addr; }
__device__ __forceinline__ double f64_of(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ uint32_t lds_u32_at(uint32_t addr) { return *(lds_u32*)(uintptr_t)addr; }
__device__ __forceinline__ void lds_store_u32(uint32_t addr, uint32_t v) { *(lds_u32*)(uintptr_t)addr = v; }
__device__ __forceinline__ void lds_add_u32(uint32_t addr, uint32_t v) {
    __hip_atomic_fetch_add((lds_u32*)(uintptr_t)addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

struct TunedWindow {
    int32_t wmin;       // window minimum (integer path)
    int32_t span;       // SIMPLE: emax - emin (= bins x bin size); else phi_hi - phi_lo
    int32_t phi_lo;
    int32_t sh;         // log2(bin size)
};

// One attempt at a site whose Delta E has been reduced to the proposed window
// offset `on`.  SIMPLE windows: bins [-guard, nbins + guard) exist in LDS, the
// guard bins holding ln g = +inf, so a proposal outside the window needs no
// range test (|Delta E| <= guard x bin size).  Otherwise the reference's
// phi(e) = (int)(e + 0.5) range test picks the +inf sentinel slot.
template <bool SIMPLE>
__device__ __forceinline__ uint32_t tuned_slot(int on, const TunedWindow& tw, int stride16, uint32_t zero_off,
                                               uint32_t sent) {
    if (SIMPLE) {
        // |Delta E| <= guard bins: in range because the walker is inside its
        // window (the host refuses to sweep before rewl_adjust_to_range)
        return (uint32_t)__mul24(on >> tw.sh, stride16) + zero_off;
    }
    const int En = on + tw.wmin;
    const bool in = (unsigned)(phi_int(En) - tw.phi_lo) < (unsigned)tw.span;
    return in ? (uint32_t)__mul24(max(on, 0) >> tw.sh, stride16) + zero_off : sent;
}

template <bool SIMPLE>
__device__ __forceinline__ void tuned_attempt(int on, const TunedWindow& tw, int& o, uint32_t& cbo, double lnf,
                                              int stride16, uint32_t zero_off, uint32_t sent, double B, int& a) {
    const uint32_t nbo = tuned_slot<SIMPLE>(on, tw, stride16, zero_off, sent);
    // the current bin is re-read rather than carried: its ln g changes only
    // through the lds_add_f64 below, which LDS orders before this read
    const double ln = lds_f64(nbo);
    const double lc = lds_f64(cbo);
    const bool acc = (ln - lc) < B;
    o = acc ? on : o;
    cbo = acc ? nbo : cbo;
    lds_add_f64(cbo, lnf);
    lds_add_u32(cbo + 8, 1u);
    a = acc ? 1 : 0;
}

// The Fourier sums f of one bit plane after the sites in `fw` flipped in row
// y (old row `old`, bit 1 = -1): a site going from -1 to +1 raises its row
// and column sums by 2, one going the other way lowers them by 2.  cs is the
// LDS address of cos(q k), k < L, followed by sin(q k).
template <int L>
__device__ __forceinline__ void fourier_flip(double f[4], uint64_t old, uint64_t fw, int y, uint32_t cs) {
    const uint64_t up = old & fw, dn = ~old & fw;
    const double dR = (double)(2 * ((int)__popcll(up) - (int)__popcll(dn)));
    f[0] = fma(dR, lds_f64(cs + 8u * (uint32_t)y), f[0]);
    f[1] = fma(dR, lds_f64(cs + 8u * (uint32_t)(L + y)), f[1]);
    uint64_t m = fw;
    while (m) {
        const int x = __builtin_ctzll(m);
        m &= m - 1ull;
        const double d = ((old >> x) & 1ull) ? 2.0 : -2.0;
        f[2] = fma(d, lds_f64(cs + 8u * (uint32_t)x), f[2]);
        f[3] = fma(d, lds_f64(cs + 8u * (uint32_t)(L + x)), f[3]);
    }
}

// OBS (with SPEC 1 only): the observables of update_observables at every
// attempt, from magnetisations kept in registers; the moments of consecutive
// attempts in one bin are summed in registers and added to the bin's sums
// when the walker leaves it.  OBS = 2 adds the correlators G(qmin).
template <int MODEL, int L, bool SIMPLE, int SPEC, int OBS = 0>
__device__ __forceinline__ void sweep_walker_tuned(const KArgs& a, const DevGroup& G, const DevSystem& S, int lane,
                                                   int nsweeps, int guard, uint32_t ring_lane, uint32_t cs_tab) {
    static_assert(!OBS || SPEC == 1, "observables are tracked by the SPEC 1 form");
    constexpr int NW = (L + 31) / 32;    // 32-bit words per row
    constexpr int NW2 = (L + 15) / 16;   // words of 2-bit fields per row
    using Mdl = Model<MODEL>;
    constexpr int NP = Mdl::kPlanes;
    constexpr int NPASS = Mdl::kPasses;
    constexpr bool KQ = Mdl::kProductCoupling;
    const bool active = lane < G.nwalkers;
    const int g = G.first_walker + (active ? lane : 0);
    const int wi = a.w_window[g];
    const DevWindow Wn = a.win[wi];
    const double lnf = a.win_lnf[wi];
    const int64_t hb = G.hist_base + lane;
    const int gbins = G.nbins;     // bins of the widest window of the group (LDS rows)
    const int nbins = Wn.nbins;    // bins of this walker's window
    const int stride = G.nwalkers;
    const int stride16 = stride * kTunedSlot;
    // slot of bin b: (b + guard) * stride16 + lane * 16, b in [-guard, nbins + guard)
    const uint32_t zero_off = lds_base() + (uint32_t)(guard * stride16 + lane * kTunedSlot);
    const uint32_t sent = zero_off + (uint32_t)(nbins * stride16);   // first upper guard bin
    // observables: the per-bin moment sums live in LDS after the ring (and
    // after the correlator's cos/sin table with observables = 2),
    // [bin][walker][kValues] FP64, added at every attempt and added to HBM
    // after the launch; the current slot's sums are at
    // obs_lane + (cbo - zero_off) / 2 * kValues
    constexpr bool LOBS = OBS >= 1;
    constexpr int NVL = Mdl::kValues;
    const uint32_t obs_lane = cs_tab + (uint32_t)(OBS == 2 ? 16 * L : 0) + (uint32_t)(lane * 8 * NVL);

    if (active) {
        for (int b = -guard; b < gbins + guard; ++b) {
            const uint32_t off = zero_off + (uint32_t)(b * stride16);
            const bool inside = b >= 0 && b < nbins;   // a narrower window's spare bins are guards too
            const uint32_t c = inside ? a.cnt[hb + (int64_t)b * kLanes] : 0u;
            lds_store_f64(off, inside ? a.lng[hb + (int64_t)b * kLanes] : __builtin_inf());
            lds_store_u32(off + 8, c);
            // all-time updates of the bin = this word + its current count
            if (OBS == 2) lds_store_u32(off + 12, inside ? a.obs_tot[hb + (int64_t)b * kLanes] - c : 0u);
        }
        if (LOBS)
            for (int b = 0; b < gbins; ++b)
#pragma unroll
                for (int i = 0; i < NVL; ++i) lds_store_f64(obs_lane + (uint32_t)((b * stride * NVL + i) * 8), 0.);
    }
    __syncthreads();   // histograms loaded, first row of thresholds ready

    TunedWindow tw;
    tw.wmin = Wn.imin;
    tw.sh = Wn.shift;
    tw.phi_lo = Wn.phi_lo;
    tw.span = SIMPLE ? (int)(Wn.emax - Wn.emin) : Wn.phi_hi - Wn.phi_lo;


    uint64_t* lat = a.lat + G.lat_base + lane;
    int fl = a.w_flavor[g];
    int o = (int)a.w_E[g] - tw.wmin;
    // the walker is inside its window (rewl_adjust_to_range); the clamp keeps
    // the slot in bounds regardless
    uint32_t cbo = (uint32_t)(min(max(o, 0) >> tw.sh, max(nbins - 1, 0)) * stride16) + zero_off;
    // ln g of the current bin, carried in a register by the read-ahead form
    double lc = active ? lds_f64(cbo) : 0.0;

    const int m4J = -2 * S.twoJ, m4K = -2 * S.twoK;
    int buf = 0;

    constexpr int NV = Mdl::kValues;
    constexpr int GK = OBS == 2 ? 4 : 3;   // slot stride of the (|x|, x^2, x^4[, G]) groups
    int mag[3] = {0, 0, 0};
    double accv[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) accv[i] = 0.;
    uint32_t acck = 0xffffffffu;   // slot of the bin being summed (none yet)
    auto flush_obs = [&]() {
        if (acck == 0xffffffffu) return;
        const int b = (int)((acck - zero_off) / (uint32_t)stride16);
        double* p = a.obs + (hb + (int64_t)b * kLanes) * a.nobs;
#pragma unroll
        for (int i = 0; i < NV; ++i) obs_add(p + (i / 3) * GK + i % 3, accv[i]);
    };
    // correlator: Fourier sums of plane 0 and, for a model with a Baxter
    // correlator, of the product of the first and last planes
    double fs[4] = {0., 0., 0., 0.}, fb[4] = {0., 0., 0., 0.};
    if (OBS == 2 && active) {
        const uint64_t* lt = a.lat + G.lat_base + lane;
        auto csf = [&](int k) { return lds_f64(cs_tab + 8u * (uint32_t)k); };
        fourier_parts(L, [&](int r) { return lt[(int64_t)r * kLanes]; }, csf, fs);
        if (Mdl::kBaxterCorrelator)
            fourier_parts(L, [&](int r) { return lt[(int64_t)r * kLanes] ^ lt[(int64_t)((NP - 1) * L + r) * kLanes]; },
                          csf, fb);
    }
    auto observe = [&]() {
        if constexpr (LOBS) {
            double v[NVL];
            Mdl::values(L * L, mag, v);
            const uint32_t oa = obs_lane + (uint32_t)__mul24((int)((cbo - zero_off) >> 1), NVL);
#pragma unroll
            for (int i = 0; i < NVL; ++i) lds_add_f64(oa + 8u * (uint32_t)i, v[i]);
            return;
        }
        if (cbo != acck) {
            flush_obs();
            acck = cbo;
#pragma unroll
            for (int i = 0; i < NV; ++i) accv[i] = 0.;
        }
        double v[NV];
        Mdl::values(L * L, mag, v);
#pragma unroll
        for (int i = 0; i < NV; ++i) accv[i] += v[i];
    };
    if (OBS && active) lattice_magnetisations<MODEL>(a.lat + G.lat_base + lane, L, mag);

    for (int s = 0; s < nsweeps; ++s) {
        for (int pass = 0; pass < NPASS; ++pass) {
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
                    // P: the plane this pass flips, Q: the other one (the
                    // product plane's partner); a one-plane model has Q = P
                    const int pf = Mdl::flip_plane(fl), qf = NP - 1 - pf;
                    const uint64_t P = plane_pick(cu, pf), Pu = plane_pick(up, pf), Pd = plane_pick(dn, pf);
                    const uint64_t Q = plane_pick(cu, qf), Qu = plane_pick(up, qf), Qd = plane_pick(dn, qf);
                    // ---- per-row static bit slices (the model's) --------------------
                    const auto sl = Mdl::template slices<L>(P, Pu, Pd, Q, Qu, Qd, rotr1<L>(P), rotl1<L>(P),
                                                            rotr1<L>(Q), rotl1<L>(Q));
                    // signed 2-bit counts, site x at bits 2 (x % 16) of word x / 16:
                    // (c0, c1n) interleaved is c0 + 2 c1 - 2 in two's complement
                    uint32_t w_cp[NW2], w_cq[NW2], w_pdl[NW], w_ql[NW], frev[NW];
#pragma unroll
                    for (int w = 0; w < NW2; ++w) {
                        w_cp[w] = spread16((uint32_t)(sl.c0 >> (16 * w)) & 0xffffu) |
                                  (spread16((uint32_t)(sl.c1n >> (16 * w)) & 0xffffu) << 1);
                        if constexpr (KQ)
                            w_cq[w] = spread16((uint32_t)(sl.m0 >> (16 * w)) & 0xffffu) |
                                      (spread16((uint32_t)(sl.m1n >> (16 * w)) & 0xffffu) << 1);
                    }
#pragma unroll
                    for (int h = 0; h < NW; ++h) {
                        w_pdl[h] = (uint32_t)(sl.pdl >> (32 * h));
                        if constexpr (KQ) w_ql[h] = (uint32_t)(sl.ql >> (32 * h));
                        frev[h] = 0u;
                    }
                    const uint32_t rrow = ring_lane + (uint32_t)(buf * 8 * L);
                    int acc = 0, f0 = 0;
                    // bounds of attempts (x, x + 1) in bnd, of (x + 2, x + 3) already in flight
                    f64x2 bnd = lds_f64x2(rrow), bnd_next = lds_f64x2(rrow + 16u);
                    // proposal offset of site x when the previous site's outcome is `a`
                    auto propose = [&](int x, int a, int from) -> int {
                        const int h = x >> 5, b = x & 31;
                        const uint32_t dl = bit_of(w_pdl[h], b) ^ (uint32_t)a;
                        // Delta E = -4J (cP' + dl) [ - 4K (cQ' + (dl ^ ql)) ]
                        int kp = (int)dl + __builtin_amdgcn_sbfe((int)w_cp[x >> 4], 2 * (x & 15), 2);
                        int kq = 0;
                        if constexpr (KQ)
                            kq = (int)(dl ^ bit_of(w_ql[h], b)) + __builtin_amdgcn_sbfe((int)w_cq[x >> 4], 2 * (x & 15), 2);
                        if (x == L - 1) {
                            // the right neighbour is site 0 of this row: its static
                            // disagreement toggles if it flipped
                            const int prl = (int)bit_of((uint32_t)(sl.pr >> (32 * h)), b);
                            kp += f0 * (1 - 2 * prl);
                            if constexpr (KQ) {
                                const int xrl = (int)bit_of((uint32_t)(sl.xr >> (32 * h)), b);
                                kq += f0 * (1 - 2 * xrl);
                            }
                        }
                        int on = __mul24(m4J, kp) + from;
                        if constexpr (KQ) on += __mul24(m4K, kq);
                        return on;
                    };
                    if (SPEC == 2) {
                        // Read-ahead: for both outcomes of the current attempt, the next
                        // proposal's slot is computed and its ln g read while the
                        // current decision is made; the next decision selects.  A
                        // read issued before this attempt's update is stale exactly
                        // when it hit the bin just updated, whose value is lc.
                        int onA = propose(0, 0, o), onB = onA;
                        uint32_t nboA = tuned_slot<SIMPLE>(onA, tw, stride16, zero_off, sent), nboB = nboA;
                        double l0 = lds_f64(nboA), l1 = l0;
                        bool pa = false;
#pragma unroll
                        for (int x = 0; x < L; ++x) {
                            // keep each attempt's reads ahead of the next attempt's use:
                            // no instruction moves across this point
                            __builtin_amdgcn_sched_barrier(0);
                            const int h = x >> 5;
                            if ((x & 1) == 0 && x > 0) {
                                bnd = bnd_next;
                                if (x + 2 < L) bnd_next = lds_f64x2(rrow + 8u * (uint32_t)(x + 2));
                            }
                            const double Bx = (x & 1) ? bnd.y : bnd.x;
                            const int on = pa ? onB : onA;
                            const uint32_t nbo = pa ? nboB : nboA;
                            int on0 = 0, on1 = 0;
                            uint32_t nbo0 = 0, nbo1 = 0;
                            double r0 = 0.0, r1 = 0.0;
                            if (x + 1 < L) {
                                on0 = propose(x + 1, 0, o);
                                on1 = propose(x + 1, 1, on);
                                nbo0 = tuned_slot<SIMPLE>(on0, tw, stride16, zero_off, sent);
                                nbo1 = tuned_slot<SIMPLE>(on1, tw, stride16, zero_off, sent);
                                r0 = lds_f64(nbo0);
                                r1 = lds_f64(nbo1);
                            }
                            __builtin_amdgcn_sched_barrier(0);   // the reads issue before the decision
                            double ln = pa ? l1 : l0;
                            ln = nbo == cbo ? lc : ln;
                            const bool ac = (ln - lc) < Bx;
                            o = ac ? on : o;
                            cbo = ac ? nbo : cbo;
                            lc = (ac ? ln : lc) + lnf;
                            lds_store_f64(cbo, lc);
                            lds_add_u32(cbo + 8, 1u);
                            acc = ac ? 1 : 0;
                            if (x == 0) f0 = acc;
                            frev[h] = (frev[h] << 1) | (uint32_t)acc;
                            onA = on0; onB = on1; nboA = nbo0; nboB = nbo1; l0 = r0; l1 = r1; pa = ac;
                        }
                    } else if (SPEC == 1) {
                        // The slot of the next proposal is computed for both outcomes of
                        // the current attempt while its LDS reads are in flight; the
                        // outcome only selects, so the chain from one decision to the
                        // next read is one v_cndmask.
                        int on = propose(0, 0, o);
                        uint32_t nbo = tuned_slot<SIMPLE>(on, tw, stride16, zero_off, sent);
#pragma unroll
                        for (int x = 0; x < L; ++x) {
                            const int h = x >> 5;
                            if ((x & 1) == 0 && x > 0) {
                                bnd = bnd_next;
                                if (x + 2 < L) bnd_next = lds_f64x2(rrow + 8u * (uint32_t)(x + 2));
                            }
                            const double Bx = (x & 1) ? bnd.y : bnd.x;
                            double ln, lc;
                            u32x4v vn, vc;
                            if constexpr (OBS == 2) {
                                // ln g, count and update-count word of both candidate bins
                                vn = lds_u32x4(nbo);
                                vc = lds_u32x4(cbo);
                                ln = f64_of(vn.x, vn.y);
                                lc = f64_of(vc.x, vc.y);
                            } else {
                                ln = lds_f64(nbo);
                                lc = lds_f64(cbo);
                            }
                            int on0 = 0, on1 = 0;
                            uint32_t nbo0 = 0, nbo1 = 0;
                            if (x + 1 < L) {
                                on0 = propose(x + 1, 0, o);
                                on1 = propose(x + 1, 1, on);
                                nbo0 = tuned_slot<SIMPLE>(on0, tw, stride16, zero_off, sent);
                                nbo1 = tuned_slot<SIMPLE>(on1, tw, stride16, zero_off, sent);
                                // keep both slots materialised: selecting the offset first
                                // and deriving the slot after the decision would put the
                                // shift and multiply-add back on the chain
                                asm volatile("" : "+v"(nbo0), "+v"(nbo1));
                            }
                            const bool ac = (ln - lc) < Bx;
                            o = ac ? on : o;
                            cbo = ac ? nbo : cbo;
                            lds_add_f64(cbo, lnf);
                            lds_add_u32(cbo + 8, 1u);
                            acc = ac ? 1 : 0;
                            if (x == 0) f0 = acc;
                            frev[h] = (frev[h] << 1) | (uint32_t)acc;
                            on = ac ? on1 : on0;
                            nbo = ac ? nbo1 : nbo0;
                            if constexpr (OBS) {
                                // the flipped spin's bit in its plane and its product
                                // bit (1 = -1 before the flip)
                                Mdl::magnetise(mag, ac, (int)((P >> x) & 1ull), (int)(((P ^ Q) >> x) & 1ull), fl);
                                observe();
                            }
                            if constexpr (OBS == 2) {
                                // the bin's updates before this one (ising2d_hamiltonian.hpp:247-252:
                                // a sample at counts 0, 1000, 2000, ...)
                                const uint32_t t = ac ? vn.z + vn.w : vc.z + vc.w;
                                if (t % 1000u == 0u) {
                                    // this row's flips so far (sites 0..x), then the sums of the
                                    // current configuration
                                    uint64_t fx = 0;
#pragma unroll
                                    for (int hh = 0; hh <= (x >> 5); ++hh) {
                                        const int nh = hh < (x >> 5) ? 32 : (x & 31) + 1;
                                        fx |= (uint64_t)(__builtin_bitreverse32(frev[hh]) >> (32 - nh)) << (32 * hh);
                                    }
                                    double gs[4] = {fs[0], fs[1], fs[2], fs[3]};
                                    double gb[4] = {fb[0], fb[1], fb[2], fb[3]};
                                    if (pf == 0) fourier_flip<L>(gs, P, fx, y, cs_tab);
                                    if (Mdl::kBaxterCorrelator) fourier_flip<L>(gb, P ^ Q, fx, y, cs_tab);
                                    const int b = (int)((cbo - zero_off) / (uint32_t)stride16);
                                    double* p = a.obs + (hb + (int64_t)b * kLanes) * a.nobs;
                                    Mdl::add_correlators([&](int slot, double gv) { obs_add(p + slot, gv); },
                                                         fourier_g(L, gs),
                                                         Mdl::kBaxterCorrelator ? fourier_g(L, gb) : 0.);
                                }
                            }
                        }
                    } else {
#pragma unroll
                        for (int x = 0; x < L; ++x) {
                            const int h = x >> 5;
                            if ((x & 1) == 0 && x > 0) {
                                bnd = bnd_next;
                                if (x + 2 < L) bnd_next = lds_f64x2(rrow + 8u * (uint32_t)(x + 2));
                            }
                            const double Bx = (x & 1) ? bnd.y : bnd.x;
                            const int on = propose(x, acc, o);
                            tuned_attempt<SIMPLE>(on, tw, o, cbo, lnf, stride16, zero_off, sent, Bx, acc);
                            if (x == 0) f0 = acc;
                            frev[h] = (frev[h] << 1) | (uint32_t)acc;   // site x lands at bit (sites in word - 1 - x % 32)
                        }
                    }
                    uint64_t fw = 0;
#pragma unroll
                    for (int h = 0; h < NW; ++h) {
                        constexpr int last = L - 32 * (NW - 1);   // sites in the last word
                        const int nh = h == NW - 1 ? last : 32;
                        fw |= (uint64_t)(__builtin_bitreverse32(frev[h]) >> (32 - nh)) << (32 * h);
                    }
                    const uint64_t Pn = P ^ fw;
                    if constexpr (OBS == 2) {
                        if (pf == 0) fourier_flip<L>(fs, P, fw, y, cs_tab);
                        if (Mdl::kBaxterCorrelator) fourier_flip<L>(fb, P ^ Q, fw, y, cs_tab);
                    }
                    plane_set(cu, pf, Pn);
#pragma unroll
                    for (int q = 0; q < NP; ++q) {
                        lat[(int64_t)(q * L + y) * kLanes] = cu[q];
                        if (y == 0) r0[q] = cu[q];
                        up[q] = cu[q];
                        cu[q] = dn[q];
                        dn[q] = (y + 2 < L) ? nx[q] : r0[q];
                    }
                }
                __syncthreads();   // this row's bounds consumed, the next row's written
                buf ^= 1;
            }
            fl = Mdl::next_flavour(fl);
        }
    }

    if (active) {
        if (OBS && !LOBS) flush_obs();
        if (LOBS)
            for (int b = 0; b < nbins; ++b) {
                double* p = a.obs + (hb + (int64_t)b * kLanes) * a.nobs;
#pragma unroll
                for (int i = 0; i < NVL; ++i)
                    p[(i / 3) * GK + i % 3] += lds_f64(obs_lane + (uint32_t)((b * stride * NVL + i) * 8));
            }
        for (int b = 0; b < nbins; ++b) {
            const uint32_t off = zero_off + (uint32_t)(b * stride16);
            a.lng[hb + (int64_t)b * kLanes] = lds_f64(off);
            a.cnt[hb + (int64_t)b * kLanes] = lds_u32_at(off + 8);
            if (OBS == 2) a.obs_tot[hb + (int64_t)b * kLanes] = lds_u32_at(off + 12) + lds_u32_at(off + 8);
        }
        a.w_E[g] = (double)(o + tw.wmin);
        a.w_flavor[g] = fl;
    }
}

// Producer side: the acceptance bound B (accept_bound.hpp) of every attempt
// of one row, written into each walker's own ring row: walker j's bounds for
// attempt x of buffer `buf` sit at ring_base + j * ring_stride + buf * 8 L + 8 x.
// L is a multiple of 4, so a row starts on a Philox block boundary and each
// block fills four consecutive entries (two 16-byte stores).
template <int MODEL, int L>
__device__ __forceinline__ void fill_row_tuned(const DevGroup& G, const DevSystem& S, int wave, int nwaves, int lane,
                                               uint64_t sweep_base, int row, uint32_t ring_lane_buf,
                                               const BoundEntry* tab) {
    constexpr int NPASS = Model<MODEL>::kPasses;
    constexpr int N = L * L;
    constexpr uint64_t BLK = (uint64_t)(NPASS * N + 3) / 4;   // Philox blocks per sweep
    const uint32_t walker = (uint32_t)(G.first_walker - S.first_walker + lane);
    constexpr int per_sweep = NPASS * L;
    const int s = row / per_sweep, rem = row - s * per_sweep;
    const int pass = rem / L, y = rem - pass * L;
    const int b0 = (pass * N + y * L) >> 2;
    const uint64_t blk0 = (sweep_base + (uint64_t)s) * BLK;
    for (int j = wave; j < L / 4; j += nwaves) {
        const uint64_t bi = blk0 + (uint64_t)(b0 + j);
        const u32x4 r = philox4x32_10(u32x4{(uint32_t)bi, (uint32_t)(bi >> 32), DOMAIN_SWEEP, walker}, S.k0, S.k1);
        const uint32_t dst = ring_lane_buf + 32u * (uint32_t)j;
        f64x2 v0, v1;
        v0.x = accept_bound(r.x, tab);
        v0.y = accept_bound(r.y, tab);
        v1.x = accept_bound(r.z, tab);
        v1.y = accept_bound(r.w, tab);
        lds_store_f64x2(dst, v0);
        lds_store_f64x2(dst + 16u, v1);
    }
}

template <int MODEL, int L>
__device__ __forceinline__ void produce_bounds_tuned(const DevGroup& G, const DevSystem& S, int wave, int nwaves,
                                                     int lane, int nsweeps, uint64_t sweep_base, uint32_t ring_lane,
                                                     const BoundEntry* tab) {
    constexpr int NPASS = Model<MODEL>::kPasses;
    const int rows = nsweeps * NPASS * L;
    const bool mine = lane < G.nwalkers;
    if (mine) fill_row_tuned<MODEL, L>(G, S, wave, nwaves, lane, sweep_base, 0, ring_lane, tab);
    __syncthreads();   // pairs with the walkers' barrier after loading the histograms
    for (int r = 0; r < rows; ++r) {
        if (mine && r + 1 < rows)
            fill_row_tuned<MODEL, L>(G, S, wave, nwaves, lane, sweep_base, r + 1,
                                     ring_lane + (uint32_t)(((r + 1) & 1) * 8 * L), tab);
        __syncthreads();   // pairs with the walkers' end-of-row barrier
    }
}

// One workgroup = one walker wavefront + 1..7 producer wavefronts; groups are
// taken persistently in blockIdx order.  LDS: (nbins + 2 guard) x group size
// slots of 16 bytes, then one ring row of tuned_ring_stride(L) bytes per
// walker (two buffers of L FP64 bounds, padded by 16 bytes so that the
// walkers' 16-byte reads fall on different banks).
template <int MODEL, int L, bool SIMPLE, int SPEC, int OBS = 0>
__global__ void __launch_bounds__(512) k_sweep_tuned(KArgs a, const int32_t* groups, int ngroups, int nsweeps,
                                                     uint64_t sweep_base, int guard) {
    const int tid = threadIdx.x;
    // The walker wavefront shares its SIMD with a producer when there are more
    // than three producers; at the highest priority it issues whenever it is
    // ready and the producer fills the cycles it spends waiting on LDS.
    if (tid < kLanes) __builtin_amdgcn_s_setprio(3);
    for (int gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const DevGroup G = a.grp[groups[gi]];
        const DevSystem S = a.sys[G.sys];
        const uint32_t ring = lds_base() + (uint32_t)((G.nbins + 2 * guard) * G.nwalkers * kTunedSlot);
        const int lane = tid & (kLanes - 1);
        const uint32_t ring_lane = ring + (uint32_t)(lane * tuned_ring_stride(L));
        // after the ring: with observables = 2 the cos and sin of qmin k
        // (Transform_Kernel, fourier_correlator.hpp:14-42), visible to the walker
        // after its first barrier; with observables = 1 (Ising) the per-bin sums
        const uint32_t cs_tab = ring + (uint32_t)(G.nwalkers * tuned_ring_stride(L));
        if (OBS == 2) {
            const double qmin = 2. * acos(-1.) / (double)L;
            for (int k = tid; k < L; k += (int)blockDim.x) {
                lds_store_f64(cs_tab + 8u * (uint32_t)k, cos(qmin * (double)k));
                lds_store_f64(cs_tab + 8u * (uint32_t)(L + k), sin(qmin * (double)k));
            }
        }
        if (tid < kLanes)
            sweep_walker_tuned<MODEL, L, SIMPLE, SPEC, OBS>(a, G, S, tid, nsweeps, guard, ring_lane, cs_tab);
        else
            produce_bounds_tuned<MODEL, L>(G, S, tid / kLanes - 1, (int)(blockDim.x / kLanes) - 1, lane, nsweeps,
                                           sweep_base, ring_lane, static_cast<const BoundEntry*>(a.bound_tab));
        __syncthreads();   // the ring and LDS histograms are reused by the next group
    }
}

// B(k) for k = ks[i] (the sweep kernel's acceptance bounds, device log)
__global__ void k_selfcheck_bounds(const uint32_t* ks, uint32_t n, const BoundEntry* tab, double* out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = accept_bound(ks[i], tab);
}

}  // namespace rewl
