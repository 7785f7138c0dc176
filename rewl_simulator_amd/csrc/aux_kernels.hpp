// Set-up, replica-exchange, flatness/averaging and thermodynamics kernels.
#pragma once
#include "device.hpp"

namespace rewl {

__device__ __forceinline__ const DevSystem& sys_of(const KArgs& a, int g) { return a.sys[a.win[a.w_window[g]].sys]; }

__device__ __forceinline__ bool win_in_range(const DevWindow& W, double e) {
    const int t = trunc_half(e);
    return t >= W.phi_lo && t < W.phi_hi;
}
__device__ __forceinline__ int win_bin(const DevWindow& W, double e) {
    const int b = (int)((e - W.emin) / W.bs);
    return b < 0 ? 0 : (b >= W.nbins ? W.nbins - 1 : b);
}

// Ising2d::recalculate_state / Ashkin_Teller2d::recalculate_state
// (ising2d_hamiltonian.hpp:190-210, ashkin_teller2d_hamiltonian.hpp:193-217):
// the per-site terms of the model summed over the sites in order.
template <int M>
__device__ double recalc_energy(const DevSystem& S, const uint64_t* lat, int64_t base) {
    double E = 0.;
    for (int i = 0; i < S.N; ++i) {
        double own[Model<M>::kPlanes], nb[Model<M>::kPlanes][4];
        site_spins_mem<M>(S, lat, base, i, own, nb);
        E = Model<M>::add_site_energy(E, S.J, S.K, S.has_field, field_at(S, i), own, nb);
    }
    return E;
}

__global__ void k_recalc(KArgs a) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.num_walkers) return;
    const DevSystem& S = sys_of(a, g);
    a.w_E[g] = with_model(S.model, [&](auto m) { return recalc_energy<decltype(m)::value>(S, a.lat, a.w_lat[g]); });
}

// REWL_Walker::adjust_state_to_range (rewl_walker.hpp:71-115): one draw picks
// the starting site, then sites are scanned in order, accepting flips that move
// the energy toward the window.  Each proposal is change_state's (FP64, the
// reference's order); the flavour of a model with several advances at the
// wrap to site 0 and, with mode 0, also after site N-1 as change_state does
// (mode 1 keeps only the toggle at the wrap, so sigma and tau passes alternate).
template <int M>
__device__ void adjust_walker(const DevWindow& W, const DevSystem& S, uint64_t* lat, int64_t base, double& E,
                              int& fl, double u, uint64_t max_steps, int mode, int32_t* err) {
    uint64_t site = (uint64_t)(u * S.N);
    bool in = win_in_range(W, E);
    uint64_t steps = 0;
    while (!in) {
        if (++steps > max_steps) { atomicOr(err, 2); break; }
        ++site;
        if (site == (uint64_t)S.N) { site = 0; fl = Model<M>::next_flavour(fl); }
        const int plane = Model<M>::flip_plane(fl);
        const double En = proposal_energy_mem<M>(S, lat, base, (int)site, fl, E);
        if (mode == 0 && site == (uint64_t)(S.N - 1)) fl = Model<M>::next_flavour(fl);
        const bool low = trunc_half(E) < W.phi_lo, high = trunc_half(E) >= W.phi_hi;
        if ((low && En > E) || (!low && high && En <= E)) {
            const int x = (int)site % S.L, y = (int)site / S.L;
            lat[base + (int64_t)(plane * S.L + y) * kLanes] ^= (1ull << x);
            E = En;
        }
        in = win_in_range(W, E);
    }
}

template <bool REPLAY>
__global__ void k_adjust(KArgs a, uint64_t max_steps, int mode) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.num_walkers) return;
    const DevWindow W = a.win[a.w_window[g]];
    const DevSystem& S = a.sys[W.sys];
    double E = a.w_E[g];
    int fl = a.w_flavor[g];
    double u;
    if (REPLAY) {
        uint64_t c = a.w_cursor[g];
        if (c < a.ring_filled[g]) u = canonical_from_u64(a.ring[(int64_t)g * (int64_t)a.ring_cap + (int64_t)(c & (a.ring_cap - 1))]);
        else { u = 0.0; atomicOr(a.err, 1); }
        a.w_cursor[g] = c + 1;
    } else {
        u = canonical_from_u64(philox_u64(S.k0, S.k1, DOMAIN_INIT, (uint32_t)a.w_local[g], 0ull));
    }
    with_model(S.model, [&](auto m) {
        adjust_walker<decltype(m)::value>(W, S, a.lat, a.w_lat[g], E, fl, u, max_steps, mode, a.err);
        return 0;
    });
    a.w_E[g] = E;
    a.w_flavor[g] = fl;
}

// ---- std::shuffle restated for the exchange partner draw ------------------
// libstdc++ (GCC 11) shuffles n elements with a 64-bit engine by drawing swap
// positions in pairs from one uniform_int_distribution call, each call using
// Lemire's nearly divisionless reduction with a 128-bit product.
struct ExchStream {
    bool replay;
    const uint64_t* ring;
    uint64_t cap, filled;
    uint32_t k0, k1, walker;
    uint64_t cursor;
    int32_t* err;
    __device__ uint64_t next() {
        uint64_t c = cursor++;
        if (replay) {
            if (c < filled) return ring[c & (cap - 1)];
            atomicOr(err, 1);
            return ~0ull;   // ends any rejection loop below
        }
        return philox_u64(k0, k1, DOMAIN_EXCH, walker, c);
    }
    // uniform integer in [0, range) (Lemire 2019)
    __device__ uint64_t below(uint64_t range) {
        uint64_t x = next();
        uint64_t lo = x * range, hi = __umul64hi(x, range);
        if (lo < range) {
            const uint64_t threshold = (0ull - range) % range;
            for (int guard = 0; lo < threshold && guard < 4096; ++guard) {
                x = next();
                lo = x * range;
                hi = __umul64hi(x, range);
            }
        }
        return hi;
    }
};

__device__ void shuffle_partners(int* v, int n, ExchStream& st) {
    int i = 1;
    if (n <= 1) return;
    if ((n % 2) == 0) {
        const int j = (int)st.below(2);
        const int t = v[i]; v[i] = v[j]; v[j] = t;
        ++i;
    }
    while (i != n) {
        const uint64_t r = (uint64_t)i + 1;
        const uint64_t x = st.below(r * (r + 1));
        const int p1 = (int)(x / (r + 1)), p2 = (int)(x % (r + 1));
        int t = v[i]; v[i] = v[p1]; v[p1] = t; ++i;
        t = v[i]; v[i] = v[p2]; v[p2] = t; ++i;
    }
}

__device__ __forceinline__ ExchStream exch_stream(const KArgs& a, int g, bool replay) {
    const DevSystem& S = sys_of(a, g);
    ExchStream st;
    st.replay = replay;
    st.ring = replay ? a.ring + (int64_t)g * (int64_t)a.ring_cap : nullptr;
    st.cap = a.ring_cap;
    st.filled = replay ? a.ring_filled[g] : 0ull;
    st.k0 = S.k0; st.k1 = S.k1; st.walker = (uint32_t)a.w_local[g];
    st.cursor = a.w_cursor[g];
    st.err = a.err;
    return st;
}

constexpr int kMaxReplicas = 256;

// One replica-exchange phase for one pair of adjacent windows
// (rewl_simulation.hpp:228-346): replica 0 of the lower window shuffles the
// upper replicas into a random bijection, each pair accepts with
// p = exp(min(0, dlng_lower)) * exp(min(0, dlng_upper)) against a float
// uniform of the lower walker, accepted pairs swap energy and lattice, and
// every paired walker then increments its histogram at its current bin.
__global__ void __launch_bounds__(kLanes) k_exchange(KArgs a, const int32_t* pairs, int npairs, int replay) {
    __shared__ int shuffled[kMaxReplicas];
    __shared__ int accept[kMaxReplicas];
    __shared__ double s_cs[128];   // cos/sin(qmin k), k < L <= 64, for the correlators
    const int pi = blockIdx.x;
    if (pi >= npairs) return;
    const int lw = pairs[pi], uw = lw + 1;
    const DevWindow WL = a.win[lw], WU = a.win[uw];
    const DevSystem& S = a.sys[WL.sys];
    const int R = WL.R;
    const int tid = threadIdx.x;
    if (a.obs_corr) {
        const double qmin = 2. * acos(-1.) / (double)S.L;
        for (int k = tid; k < S.L; k += blockDim.x) {
            s_cs[k] = cos(qmin * (double)k);
            s_cs[S.L + k] = sin(qmin * (double)k);
        }
    }
    for (int k = tid; k < R; k += blockDim.x) shuffled[k] = R + k;
    __syncthreads();
    if (tid == 0) {
        ExchStream st = exch_stream(a, WL.first_walker, replay != 0);
        shuffle_partners(shuffled, R, st);
        a.w_cursor[WL.first_walker] = st.cursor;
    }
    __syncthreads();
    for (int r = tid; r < R; r += blockDim.x) {
        const int ga = WL.first_walker + r, gb = WU.first_walker + shuffled[r] - R;
        const double Ea = a.w_E[ga], Eb = a.w_E[gb];
        const int64_t ha = a.w_hist[ga], hb = a.w_hist[gb];
        double pa = 0., pb = 0.;
        if (win_in_range(WL, Eb))
            pa = exp(fmin(0., a.lng[ha + (int64_t)win_bin(WL, Ea) * kLanes] - a.lng[ha + (int64_t)win_bin(WL, Eb) * kLanes]));
        if (win_in_range(WU, Ea))
            pb = exp(fmin(0., a.lng[hb + (int64_t)win_bin(WU, Eb) * kLanes] - a.lng[hb + (int64_t)win_bin(WU, Ea) * kLanes]));
        const double p = pa * pb;
        int acc = 0;
        if (p != 0.) {
            ExchStream st = exch_stream(a, ga, replay != 0);
            const float u = (float)canonical_from_u64(st.next());
            a.w_cursor[ga] = st.cursor;
            acc = u < p;
        }
        accept[r] = acc;
        if (acc) {
            a.w_E[ga] = Eb;
            a.w_E[gb] = Ea;
            atomicAdd(a.exch_accepted, 1ull);
        }
    }
    __syncthreads();
    const int words = with_model(S.model, [](auto m) { return Model<decltype(m)::value>::kPlanes; }) * S.L;
    for (int r = 0; r < R; ++r) {
        if (!accept[r]) continue;
        const int ga = WL.first_walker + r, gb = WU.first_walker + shuffled[r] - R;
        const int64_t la = a.w_lat[ga], lb = a.w_lat[gb];
        for (int k = tid; k < words; k += blockDim.x) {
            const uint64_t t = a.lat[la + (int64_t)k * kLanes];
            a.lat[la + (int64_t)k * kLanes] = a.lat[lb + (int64_t)k * kLanes];
            a.lat[lb + (int64_t)k * kLanes] = t;
        }
    }
    __syncthreads();
    const double fl = a.win_lnf[lw], fu = a.win_lnf[uw];
    for (int r = tid; r < R; r += blockDim.x) {
        const int ga = WL.first_walker + r, gb = WU.first_walker + shuffled[r] - R;
        const int64_t ia = a.w_hist[ga] + (int64_t)win_bin(WL, a.w_E[ga]) * kLanes;
        const int64_t ib = a.w_hist[gb] + (int64_t)win_bin(WU, a.w_E[gb]) * kLanes;
        a.cnt[ia] += 1u;
        a.lng[ia] += fl;
        a.cnt[ib] += 1u;
        a.lng[ib] += fu;
        if (a.obs) {
            const int64_t hi[2] = {ia, ib};
            const int gw[2] = {ga, gb};
            for (int j = 0; j < 2; ++j) {
                const uint64_t* lt = a.lat + a.w_lat[gw[j]];
                with_model(S.model, [&](auto mt) {
                    constexpr int M = decltype(mt)::value;
                    int m[3];
                    lattice_magnetisations<M>(lt, S.L, m);
                    update_observables<M>(a.obs + hi[j] * a.nobs, S.N, m, a.obs_corr != 0);
                    if (Model<M>::kDensities && a.dens) update_density(a.dens, hi[j], S.N, m);
                    if (a.obs_corr) {
                        const uint32_t t = a.obs_tot[hi[j]];
                        a.obs_tot[hi[j]] = t + 1u;
                        if (t % 1000u == 0u) add_correlators<M>(a.obs + hi[j] * a.nobs, lt, S.L, -1, nullptr, s_cs);
                    }
                    return 0;
                });
            }
        }
    }
}

// ---- flatness, averaging, reset and ln f schedule, one wavefront per window --
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
    for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o); v = w < v ? w : v; }
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o); v = w > v ? w : v; }
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

constexpr int kMaxCheckBins = 8192;   // largest window the visited-support flatness test handles

// The sum of v(0) ... v(R-1) over a window's replicas, associated as the
// reference's MPI_Allreduce(SUM) over the window communicator associates it
// (rewl_simulation.hpp:383, 416-429) under MPICH's recursive doubling: with
// pof2 the largest power of two <= R and rem = R - pof2, the first 2 rem
// values are added in pairs, then the pof2 partial values are added as a
// balanced binary tree in rank order.  The tree is built left to right with
// a carry stack (one partial sum per level).
template <class F>
__device__ __forceinline__ double replica_sum(int R, F v) {
    int pof2 = 1, top = 0;
    while (pof2 * 2 <= R) { pof2 *= 2; ++top; }
    const int rem = R - pof2;
    double st[9];
    for (int k = 0; k < pof2; ++k) {
        double x = k < rem ? v(2 * k) + v(2 * k + 1) : v(k + rem);
        int lv = 0;
        while ((k >> lv) & 1) { x = st[lv] + x; ++lv; }
        st[lv] = x;
    }
    return st[top];
}

struct CheckParams {
    int32_t flatness_mode;
    int32_t schedule;
    int32_t reduce_logdos;
    int32_t average_mode;
    double flatness;
    double lnf_final;
    int64_t pending_sweeps;
};

// check_if_window_is_flat + average_and_redistribute_window + reset_counts +
// the ln f schedule (rewl_simulation.hpp:458-470, 361-454, 541-598;
// rewl_histograms.hpp:83-95, 115-152).
__global__ void __launch_bounds__(kLanes) k_check(KArgs a, CheckParams cp, int32_t* flat_out) {
    const int wi = blockIdx.x;
    if (wi >= a.num_windows) return;
    const int lane = threadIdx.x;
    if (!a.win_active[wi]) {   // a retired system: state kept as it was
        if (lane == 0) flat_out[wi] = 0;
        return;
    }
    const DevWindow W = a.win[wi];
    const int R = W.R, nb = W.nbins;
    const uint8_t* ex = a.exclude + W.excl_off;
    const bool oot_on = cp.schedule == 1 && a.win_oot[wi];

    // mode 3: bins that no replica of the window has ever visited are left out
    __shared__ uint8_t unseen[kMaxCheckBins];
    if (cp.flatness_mode == 3) {
        for (int b = lane; b < nb; b += kLanes) {
            bool seen = false;
            for (int r = 0; r < R && !seen; ++r) {
                const int64_t i = a.w_hist[W.first_walker + r] + (int64_t)b * kLanes;
                seen = (a.visits[i] + a.cnt[i]) != 0ull;
            }
            unseen[b] = seen ? 0 : 1;
        }
        __syncthreads();
    }
    bool flat = true;
    if (!oot_on) {
        for (int r = 0; r < R && flat; ++r) {
            const int64_t h = a.w_hist[W.first_walker + r];
            uint32_t mn = 0xFFFFFFFFu, mx = 0u;
            unsigned long long sum = 0ull;
            int ne = 0, zero = 0;
            for (int b = lane; b < nb; b += kLanes) {
                if (ex[b]) continue;
                if (cp.flatness_mode == 3 && unseen[b]) continue;
                const uint32_t c = a.cnt[h + (int64_t)b * kLanes];
                mn = c < mn ? c : mn;
                mx = c > mx ? c : mx;
                sum += c;
                ++ne;
                zero |= (c == 0u);
            }
            mn = wave_min(mn); mx = wave_max(mx); sum = wave_sum(sum); ne = wave_sum(ne); zero = wave_max(zero);
            bool f;
            if (ne == 0) f = true;
            else if (cp.flatness_mode == 0) f = !zero;
            else if (cp.flatness_mode == 1)
                f = (mn != 0u ? (static_cast<float>(mx - mn) / static_cast<float>(mn)) : 3.402823466e+38f) <= (float)cp.flatness;
            else f = (mn != 0u) && ((double)mn >= cp.flatness * ((double)sum / (double)ne));
            flat = flat && f;
        }
    }

    if (flat) {
        if (R > 1 && cp.average_mode == 1) {
            const double wgt = 1. / static_cast<double>(R);
            for (int b = lane; b < nb; b += kLanes) {
                const double acc = replica_sum(R, [&](int r) {
                    return wgt * a.lng[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes];
                });
                for (int r = 0; r < R; ++r) a.lng[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes] = acc;
            }
        } else if (R > 1) {
            for (int b = lane; b < nb; b += kLanes) {
                unsigned long long tot = 0ull;
                for (int r = 0; r < R; ++r) tot += a.cnt[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes];
                if (tot == 0ull) continue;
                const double acc = replica_sum(R, [&](int r) {
                    const int64_t i = a.w_hist[W.first_walker + r] + (int64_t)b * kLanes;
                    return (static_cast<double>(a.cnt[i]) / static_cast<double>(tot)) * a.lng[i];
                });
                for (int r = 0; r < R; ++r) a.lng[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes] = acc;
            }
        }
        __syncthreads();
        for (int r = 0; r < R; ++r) {
            const int64_t h = a.w_hist[W.first_walker + r];
            const double reducer = -a.lng[h] + 0.0;
            __syncthreads();
            for (int b = lane; b < nb; b += kLanes) {
                const int64_t i = h + (int64_t)b * kLanes;
                a.visits[i] += a.cnt[i];
                a.cnt[i] = 0u;
                if (cp.reduce_logdos) a.lng[i] += reducer;
            }
            __syncthreads();
        }
    }
    if (lane == 0) {
        int64_t sw = a.win_sweeps[wi] + cp.pending_sweeps;
        if (flat) {
            int64_t it = a.win_iter[wi];
            double lnf = a.win_lnf[wi];
            int oot = a.win_oot[wi];
            const double t = static_cast<double>(sw * it);
            if (cp.schedule == 1 && !oot) oot = (lnf <= 1. / t);
            const bool on = cp.schedule == 1 && oot;
            lnf = on ? (0.5 * lnf) * 0.0 + 1. / t : (0.5 * lnf);
            a.win_lnf[wi] = lnf;
            a.win_oot[wi] = oot;
            a.win_done[wi] = !(lnf >= cp.lnf_final);
            a.win_iter[wi] = it + 1;
            sw = 0;
        }
        a.win_sweeps[wi] = sw;
        flat_out[wi] = flat ? 1 : 0;
    }
}

// Final intra-window average weighted by all-time visits
// (rewl_simulation.hpp:687 with simulation_incomplete == false), or by 1/R.
__global__ void __launch_bounds__(kLanes) k_finalize(KArgs a, int average_mode) {
    const int wi = blockIdx.x;
    if (wi >= a.num_windows) return;
    const DevWindow W = a.win[wi];
    const int R = W.R;
    if (R == 1) return;
    for (int b = threadIdx.x; b < W.nbins; b += kLanes) {
        if (average_mode == 1) {
            const double wgt = 1. / static_cast<double>(R);
            const double acc = replica_sum(R, [&](int r) {
                return wgt * a.lng[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes];
            });
            for (int r = 0; r < R; ++r) a.lng[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes] = acc;
            continue;
        }
        auto updates = [&](int r) {
            const int64_t i = a.w_hist[W.first_walker + r] + (int64_t)b * kLanes;
            return (double)(a.visits[i] + a.cnt[i]);
        };
        const double tot = replica_sum(R, updates);
        if (tot == 0.) continue;
        const double acc = replica_sum(R, [&](int r) {
            return (updates(r) / tot) * a.lng[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes];
        });
        for (int r = 0; r < R; ++r) a.lng[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes] = acc;
    }
}

// ---- multicanonical refinement (no counterpart in the reference) -------------
// Every (bin, walker) histogram entry of the context: visits and counts
// cleared (a fill kernel, not a memset, so that the step can be captured), or
// counts rolled into visits.
__global__ void k_hist_clear(KArgs a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a.visits[i] = 0ull;
        a.cnt[i] = 0u;
    }
}
__global__ void k_hist_roll(KArgs a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        a.visits[i] += a.cnt[i];
        a.cnt[i] = 0u;
    }
}

// After sampling with fixed weights 1/g (ln f = 0) the histogram of a window
// is H(b) proportional to g(b) / g_est(b): ln g_est(b) + ln H(b) estimates
// ln g(b) up to the window's constant.  H sums the window's replicas; bins no
// replica visited keep their ln g.  One wavefront per window, lanes over bins;
// every replica receives the corrected curve (replica 0's ln g, which the
// preceding average made common to all replicas).
__global__ void __launch_bounds__(kLanes) k_refine(KArgs a) {
    const int wi = blockIdx.x;
    if (wi >= a.num_windows) return;
    const DevWindow W = a.win[wi];
    for (int b = threadIdx.x; b < W.nbins; b += kLanes) {
        unsigned long long H = 0ull;
        for (int r = 0; r < W.R; ++r) {
            const int64_t i = a.w_hist[W.first_walker + r] + (int64_t)b * kLanes;
            H += a.visits[i] + a.cnt[i];
        }
        if (H == 0ull) continue;
        const double v = a.lng[a.w_hist[W.first_walker] + (int64_t)b * kLanes] + log((double)H);
        for (int r = 0; r < W.R; ++r) a.lng[a.w_hist[W.first_walker + r] + (int64_t)b * kLanes] = v;
    }
}

// Thermodynamics::calculate_thermodynamics, trapezoid rule, energy part
// (thermodynamics.hpp:77-236): one thread per temperature, bins summed from
// the top bin down in FP64 as the reference does.
__global__ void k_thermo(const double* E, const double* lng, int nb, int N, double tmin, double dT, int nT,
                         double* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const double T = tmin + dT * (double)t;
    double mx = lng[nb - 1] - E[nb - 1] / T;
    for (int b = 2; b <= nb; ++b) {
        const int cb = nb - b;
        const double v = lng[cb] - E[cb] / T;
        mx = (mx < v ? v : mx);
    }
    double part = 0., U = 0., U2 = 0., Sx = 0.;
    for (int b = 1; b <= nb; ++b) {
        const int cb = nb - b;
        const double w = exp((lng[cb] - E[cb] / T) - mx);
        const bool end = (cb == 0 || cb == nb - 1);
        part += end ? w : 2. * w;
        const double ev = E[cb];
        if (end) { U += 0.5 * ev * w; U2 += 0.5 * ev * ev * w; Sx += 0.5 * lng[cb] * w; }
        else { U += ev * w; U2 += ev * ev * w; Sx += lng[cb] * w; }
    }
    part = 0.5 * part;
    double* o = out + 5 * (size_t)t;
    o[0] = -T * (mx + log(part)) / (double)N;
    U /= ((double)N * part);
    U2 /= ((double)N * (double)N * part);
    Sx /= ((double)N * part);
    o[1] = U; o[2] = U2; o[3] = Sx;
    o[4] = (double)N * (U2 - U * U) / (T * T);
}

// Canonical averages of per-bin observables (the observable part of
// Thermodynamics::calculate_thermodynamics, thermodynamics.hpp:156-236):
// <O>(T) = sum_b t_b O_b w_b / Z with w_b = exp(ln g_b - E_b/T - max), t_b = 1/2
// at the two end bins and 1 inside, Z the trapezoid partition function.
constexpr int kThermoMaxObs = 16;
__global__ void k_thermo_obs(const double* E, const double* lng, const double* obs, int nb, int nobs, double tmin,
                             double dT, int nT, double* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const double T = tmin + dT * (double)t;
    double mx = lng[nb - 1] - E[nb - 1] / T;
    for (int b = 2; b <= nb; ++b) {
        const int cb = nb - b;
        const double v = lng[cb] - E[cb] / T;
        mx = (mx < v ? v : mx);
    }
    double part = 0.;
    double acc[kThermoMaxObs];
    for (int o = 0; o < kThermoMaxObs; ++o) acc[o] = 0.;
    for (int b = 1; b <= nb; ++b) {
        const int cb = nb - b;
        const double w = exp((lng[cb] - E[cb] / T) - mx);
        const bool end = (cb == 0 || cb == nb - 1);
        part += end ? w : 2. * w;
        for (int o = 0; o < nobs; ++o) {
            const double ov = obs[(size_t)cb * nobs + o];
            acc[o] += end ? 0.5 * ov * w : ov * w;
        }
    }
    part = 0.5 * part;
    for (int o = 0; o < nobs; ++o) out[(size_t)t * nobs + o] = acc[o] / part;
}

}  // namespace rewl
