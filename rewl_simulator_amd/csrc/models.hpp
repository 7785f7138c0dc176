// Device-side model plug-ins.
//
// The reference chooses its Hamiltonian at compile time: model_hamiltonians.hpp
// aliases Hamiltonian_t / State_t / Observables_t to one model class, and the
// walker only calls that class's local_energy, change_state, set_state,
// switch_flavor_to_update and update_observables (ising2d_hamiltonian.hpp:
// 80-303, ashkin_teller2d_hamiltonian.hpp:115-318).  Here every kernel is a
// template over the model index, and everything model-specific that a kernel
// needs comes from Model<M>: the bit planes of the packed lattice, the
// flavours swept per sweep, the FP64 energy change of a proposal in the
// reference's operation order, the integer energy change from packed
// neighbour bits, and the observables.  A new Hamiltonian on the periodic
// square lattice is one more specialisation; no kernel branches on the model.
#pragma once
#include <cstdint>

namespace rewl {

template <int M>
struct Model;

// A kernel that serves walkers of several systems (set-up, exchange) turns
// the system's runtime model index into the template argument here, once per
// walker: fn(ModelTag<M>{}) with M the model's index.
template <int M>
struct ModelTag {
    static constexpr int value = M;
};
template <class Fn>
__device__ __forceinline__ auto with_model(int model, Fn fn) {
    return model == 0 ? fn(ModelTag<0>{}) : fn(ModelTag<1>{});
}

// The FP64 local-energy sums add products c * s1 * s2 ... of a coupling (or
// field) c and spins s = +-1 one by one: en += c s1 s2.  The product is exact,
// so en = fma(c s1, s2, en) rounds once exactly where en + (c s1 s2) does.
__device__ __forceinline__ double add_term(double en, double cs, double s) { return fma(cs, s, en); }

// ---- Ising2d (ising2d_hamiltonian.hpp) --------------------------------------
template <>
struct Model<0> {
    static constexpr int kPlanes = 1;    // sigma
    static constexpr int kPasses = 1;    // num_DoF / N: one flavour
    static constexpr int kValues = 3;    // |m|, m^2, m^4 per update
    static constexpr bool kDensities = false;
    // the integer energy has a second coupling 2K on the product plane
    // (sigma tau): the packed kernels count its disagreements too
    static constexpr bool kProductCoupling = false;

    // the plane a proposal at flavour fl flips, and the flavour after a pass
    __device__ static constexpr int flip_plane(int) { return 0; }
    __device__ static constexpr int next_flavour(int fl) { return fl; }

    // local_energy (ising2d_hamiltonian.hpp:162-188) of spin own[0] with
    // field h and neighbours nb[0][k] in the reference order (x,y-1), (x-1,y),
    // (x+1,y), (x,y+1): -s f, f = h + J s_k summed in that order.
    __device__ static double local_energy(double J, double, bool, double h, const double own[1],
                                          const double nb[1][4]) {
        double f = h;
#pragma unroll
        for (int k = 0; k < 4; ++k) f = add_term(f, J, nb[0][k]);
        return -1. * own[0] * f;
    }
    // recalculate_state's per-site term (ising2d_hamiltonian.hpp:190-210)
    __device__ static double add_site_energy(double E, double J, double K, bool hf, double h, const double own[1],
                                             const double nb[1][4]) {
        return E + (-0.5 * h * own[0] + 0.5 * local_energy(J, K, hf, h, own, nb));
    }
    // E' of flipping the spin (change_state, ising2d_hamiltonian.hpp:216-222):
    // E + (-1 t) f - (-1 s) f with t = -s
    __device__ static double proposal_energy(double J, double, bool, double h, double E, const double own[1],
                                             const double nb[1][4], int) {
        double f = h;
#pragma unroll
        for (int k = 0; k < 4; ++k) f = add_term(f, J, nb[0][k]);
        const double sv = own[0], t = -sv;
        return E + -1. * t * f - -1. * sv * f;
    }

    // Packed integer path.  One row's static slices: per site the count of
    // disagreeing up, down and right neighbours as the signed 2-bit field
    // (c0, c1n) = count - 2, and the left disagreement pdl; flipping site x
    // changes E by -4J (count - 2 + left) with left = pdl ^ (site x-1 flipped).
    struct Slices {
        uint64_t c0, c1n, pdl, pr;
    };
    template <int L>
    __device__ static Slices slices(uint64_t P, uint64_t Pu, uint64_t Pd, uint64_t, uint64_t, uint64_t,
                                    uint64_t rotr_P, uint64_t rotl_P, uint64_t, uint64_t) {
        Slices s;
        const uint64_t pu = P ^ Pu, pd = P ^ Pd, pr = P ^ rotr_P;
        s.c0 = pu ^ pd ^ pr;
        s.c1n = ~((pu & pd) | (pr & (pu ^ pd)));
        s.pdl = P ^ rotl_P;
        s.pr = pr;
        return s;
    }

    // update_observables (ising2d_hamiltonian.hpp:234-256): |m|, m^2, m^4 of
    // m = M / N
    __device__ static void values(int N, const int m[3], double* v) {
        const double divide_N = 1. / (double)N;
        const double x = divide_N * fabs((double)m[0]);
        v[0] = x; v[1] = x * x; v[2] = x * x * x * x;
    }
    // magnetisation change of an update that flipped (acc) the spin whose
    // bit (1 = -1) before the flip was pb; qb is the other plane's bit
    __device__ static void magnetise(int m[3], bool acc, int pb, int, int) { m[0] += acc ? 4 * pb - 2 : 0; }

    // Integer energy change of the proposal from the bits (1 = -1) of the
    // site and of its neighbours in the reference order, for integer
    // couplings 2J (twoJ): 2J (4 - 2 a), a = disagreeing neighbours.
    __device__ static int int_energy_change(int twoJ, int, const uint32_t own[1], const uint32_t nb[1][4], int) {
        const uint32_t na = (own[0] ^ nb[0][0]) + (own[0] ^ nb[0][1]) + (own[0] ^ nb[0][2]) + (own[0] ^ nb[0][3]);
        return twoJ * (4 - 2 * (int)na);
    }

    // Host side, integer couplings 2J: the spacing of the energy levels and
    // the lowest energy (all spins aligned), -N 2J
    static long long energy_step(long long twoJ, long long) { return 2 * twoJ; }
    static long long ground_energy(long long N, long long twoJ, long long) { return -N * twoJ; }

    // Correlator samples G of the sigma field (update_qmin_correlator,
    // ising2d_hamiltonian.hpp:247-252): slot 3 of the bin's group
    static constexpr bool kBaxterCorrelator = false;
    template <class AddFn>
    __device__ static void add_correlators(AddFn add, double g_sigma, double) { add(3, g_sigma); }
};

// ---- Ashkin_Teller2d (ashkin_teller2d_hamiltonian.hpp) -----------------------
template <>
struct Model<1> {
    static constexpr int kPlanes = 2;    // sigma, tau
    static constexpr int kPasses = 2;    // sigma then tau (switch_flavor_to_update at site N - 1)
    static constexpr int kValues = 12;   // sigma, tau, phi, Baxter: each |x|, x^2, x^4
    static constexpr bool kDensities = true;
    static constexpr bool kProductCoupling = true;

    __device__ static constexpr int flip_plane(int fl) { return fl; }
    __device__ static constexpr int next_flavour(int fl) { return fl ^ 1; }

    // local_energy (ashkin_teller2d_hamiltonian.hpp:169-191) of the spins
    // (sigma, tau) = own: -en, en = h (s t) + sum_k [J s s_k + J t t_k +
    // K s t s_k t_k] in the reference order (field term only with a field).
    __device__ static double local_energy(double J, double K, bool has_field, double h, const double own[2],
                                          const double nb[2][4]) {
        const double sg = own[0], ta = own[1];
        const double Js = J * sg, Jt = J * ta, Kst = K * sg * ta;
        double en = 0.;
        if (has_field) en += h * (sg * ta);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            en = add_term(en, Js, nb[0][k]);
            en = add_term(en, Jt, nb[1][k]);
            en = add_term(en, Kst, nb[0][k] * nb[1][k]);
        }
        return -en;
    }
    // recalculate_state's per-site terms (ashkin_teller2d_hamiltonian.hpp:193-217)
    __device__ static double add_site_energy(double E, double J, double K, bool has_field, double h,
                                             const double own[2], const double nb[2][4]) {
        if (has_field) E += -0.5 * h * own[0] * own[1];
        return E + 0.5 * local_energy(J, K, has_field, h, own, nb);
    }
    // E' of flipping sigma (fl = 0) or tau (fl = 1) of a site (change_state,
    // ashkin_teller2d_hamiltonian.hpp:223-250): E + local(new) - local(old),
    // each sum in the reference's order.  One of sigma, tau flips, so the
    // proposal's products K s t and h (s t) are the current ones negated,
    // exactly (every factor is +-1 or the coupling, a negation is exact).
    __device__ static double proposal_energy(double J, double K, bool has_field, double h, double E,
                                             const double own[2], const double nb[2][4], int fl) {
        const double sg = own[0], ta = own[1];
        const double Js = J * sg, Jt = J * ta, Kst = K * sg * ta;
        const double hst = has_field ? h * (sg * ta) : 0.;
        const double Jsn = fl == 0 ? -Js : Js, Jtn = fl == 1 ? -Jt : Jt;
        double en_new = 0., en_old = 0.;
        if (has_field) { en_new += -hst; en_old += hst; }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double nn = nb[0][k] * nb[1][k];
            en_new = add_term(en_new, Jsn, nb[0][k]);
            en_new = add_term(en_new, Jtn, nb[1][k]);
            en_new = add_term(en_new, -Kst, nn);
            en_old = add_term(en_old, Js, nb[0][k]);
            en_old = add_term(en_old, Jt, nb[1][k]);
            en_old = add_term(en_old, Kst, nn);
        }
        return E + -en_new - -en_old;
    }

    // Packed integer path: the flipped plane P (and the other plane Q) give
    // the J part as for Ising, and the mixed products (P ^ Q agreement, the
    // K part) a second signed 2-bit count (m0, m1n) with its left bit ql.
    struct Slices {
        uint64_t c0, c1n, pdl, pr, m0, m1n, ql, xr;
    };
    template <int L>
    __device__ static Slices slices(uint64_t P, uint64_t Pu, uint64_t Pd, uint64_t Q, uint64_t Qu, uint64_t Qd,
                                    uint64_t rotr_P, uint64_t rotl_P, uint64_t rotr_Q, uint64_t rotl_Q) {
        Slices s;
        const uint64_t pu = P ^ Pu, pd = P ^ Pd, pr = P ^ rotr_P;
        s.c0 = pu ^ pd ^ pr;
        s.c1n = ~((pu & pd) | (pr & (pu ^ pd)));
        s.pdl = P ^ rotl_P;
        s.pr = pr;
        const uint64_t qu = Q ^ Qu, qd = Q ^ Qd, qr = Q ^ rotr_Q;
        const uint64_t xu = pu ^ qu, xd = pd ^ qd;
        s.xr = pr ^ qr;
        s.m0 = xu ^ xd ^ s.xr;
        s.m1n = ~((xu & xd) | (s.xr & (xu ^ xd)));
        s.ql = Q ^ rotl_Q;
        return s;
    }

    // update_observables (ashkin_teller2d_hamiltonian.hpp:267-318): sigma, tau,
    // phi = (sigma, tau) as a vector, Baxter sigma tau
    __device__ static void values(int N, const int m[3], double* v) {
        const double divide_N = 1. / (double)N;
        const double sv = divide_N * fabs((double)m[0]), tv = divide_N * fabs((double)m[1]);
        const double bv = divide_N * fabs((double)m[2]);
        const double phi2 = sv * sv + tv * tv;
        v[0] = sv; v[1] = sv * sv; v[2] = sv * sv * sv * sv;
        v[3] = tv; v[4] = tv * tv; v[5] = tv * tv * tv * tv;
        v[6] = sqrt(phi2); v[7] = phi2; v[8] = phi2 * phi2;
        v[9] = bv; v[10] = bv * bv; v[11] = bv * bv * bv * bv;
    }
    // pb: the flipped spin's bit before the flip (plane fl); xb: its Baxter bit
    // (sigma ^ tau) before the flip
    __device__ static void magnetise(int m[3], bool acc, int pb, int xb, int fl) {
        const int d = acc ? 4 * pb - 2 : 0;
        m[0] += fl ? 0 : d;
        m[1] += fl ? d : 0;
        m[2] += acc ? 4 * xb - 2 : 0;
    }

    // 2J (4 - 2 a_P) + 2K (4 - 2 a_X): a_P disagreeing neighbours of the
    // flipped plane, a_X of the products sigma tau
    __device__ static int int_energy_change(int twoJ, int twoK, const uint32_t own[2], const uint32_t nb[2][4],
                                            int fl) {
        const uint32_t ms = (own[0] ^ nb[0][0]) | ((own[0] ^ nb[0][1]) << 1) | ((own[0] ^ nb[0][2]) << 2) |
                            ((own[0] ^ nb[0][3]) << 3);
        const uint32_t mt = (own[1] ^ nb[1][0]) | ((own[1] ^ nb[1][1]) << 1) | ((own[1] ^ nb[1][2]) << 2) |
                            ((own[1] ^ nb[1][3]) << 3);
        const int pf = __popc(fl ? mt : ms);
        const int qq = __popc(ms ^ mt);
        return twoJ * (4 - 2 * pf) + twoK * (4 - 2 * qq);
    }

    // Host side: level spacing 2 gcd(2J, 2K) and the energy of the all-up
    // configuration, -N (2 2J + 2K)
    static long long energy_step(long long twoJ, long long twoK) {
        long long a = twoJ < 0 ? -twoJ : twoJ, b = twoK < 0 ? -twoK : twoK;
        while (b) { const long long t = a % b; a = b; b = t; }
        return 2 * a;
    }
    static long long ground_energy(long long N, long long twoJ, long long twoK) { return -N * (2 * twoJ + twoK); }

    // Correlator samples (ashkin_teller2d_hamiltonian.hpp:296-316): the tau
    // correlator repeats the sigma one and phi is their sum; Baxter from the
    // product field sigma tau.  Slots 3, 7, 11, 15.
    static constexpr bool kBaxterCorrelator = true;
    template <class AddFn>
    __device__ static void add_correlators(AddFn add, double g_sigma, double g_baxter) {
        add(3, g_sigma);
        add(7, g_sigma);
        add(11, g_sigma + g_sigma);
        add(15, g_baxter);
    }
};

// Host side: the constants of the model with runtime index m (the C ABI's
// REWL_ISING2D = 0, REWL_ASHKIN_TELLER2D = 1), from the specialisations above.
struct ModelInfo {
    int planes, passes, values;
    bool product_coupling;
    long long (*energy_step)(long long, long long);
    long long (*ground_energy)(long long, long long, long long);
};
template <int M>
constexpr ModelInfo model_info_of() {
    return ModelInfo{Model<M>::kPlanes, Model<M>::kPasses, Model<M>::kValues, Model<M>::kProductCoupling,
                     &Model<M>::energy_step, &Model<M>::ground_energy};
}
inline ModelInfo model_info(int m) { return m == 0 ? model_info_of<0>() : model_info_of<1>(); }

}  // namespace rewl
