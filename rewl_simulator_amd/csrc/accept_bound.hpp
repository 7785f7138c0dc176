// Wang-Landau acceptance as one FP64 comparison.
//
// The reference accepts a move with d = (float)(ln g[new] - ln g[cur]) when
// d <= 0, and otherwise when u < exp(-(double)d) for its uniform draw u
// (wang_landau.hpp:72-84).  For a draw u = k 2^-32 (k a 32-bit word), with
// exp correctly rounded, u < RN(exp(-d)) <=> exp(-d) > u + h, h = half an
// FP64 ulp of u, <=> d < T = -ln(u + h) = -ln u - h/u (to far below an ulp
// of T).  The h/u term matters for u near 1, where T ~ 2^-32 and floats are
// finer than the FP64 grid of exp near 1.  Since d is a float, d < T is
// d <= f with f the largest float below T; d <= 0 is contained (f > 0).
// For k = 0 (u = 0) the move is accepted while exp(-d) > 0, i.e. for
// d <= 0x1.74910cp+9 = 745.133179..., the largest float whose FP64 exp(-d)
// does not underflow to zero (ln 2^-1075 = -745.13321910...).
//
// Rounding x = ln g[new] - ln g[cur] to float gives a value <= f exactly when
// x < m, m = the midpoint between f and the next float f+, or x == m and f is
// even (ties to even).  So the whole rule is x < B with
//     B = m           if f is odd,
//     B = next(m)     if f is even (the next double above m),
// a single FP64 comparison in the walker; B is computed off the walker's
// critical path.  The rule reproduces a correctly rounded exp; a libm exp
// within 0.502 ulp (glibc's bound) can differ from it only where exp(-d) lies
// within 0.002 ulp of the rounding midpoint u + h.
//
// T is evaluated as C - ln(1 + x) with a 1537-entry table (the same table on
// the host and the device, built once from the host's log1p) and an order-8
// series, |x| < 2^-9, in explicit FMAs:
//   u >= 1/2:  u0 = 1 - i 2^-10 >= u (i = top bits of 2^32 - k), C = -ln u0,
//              x = (u - u0)/u0 <= 0, so T = C + |ln(1 + x)| has no
//              cancellation, down to u0 = 1, C = 0 for the top 2^22 words;
//   u <  1/2:  k = 2^e m, m0 = 1 + i 2^-10 <= m, C = (32 - e) ln 2 - ln m0
//              (ln 2 split hi + lo, the hi part exact in the product), T >= ln 2.
// T is within a few FP64 ulps; that moves f only where a float lies that close
// to T (about 2^-28 of the words), and there only d == f decides differently.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define REWL_AB_HD __host__ __device__ __forceinline__
#else
#define REWL_AB_HD inline
#endif

namespace rewl {

REWL_AB_HD uint32_t float_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof(u));
    return u;
}
REWL_AB_HD float bits_float(uint32_t u) {
    float f;
    memcpy(&f, &u, sizeof(f));
    return f;
}
REWL_AB_HD uint64_t double_bits(double d) {
    uint64_t u;
    memcpy(&u, &d, sizeof(u));
    return u;
}
REWL_AB_HD double bits_double(uint64_t u) {
    double d;
    memcpy(&d, &u, sizeof(d));
    return d;
}

// B for a positive float f: x < B <=> (float)x <= f.
REWL_AB_HD double float_le_bound(float f) {
    const uint32_t fb = float_bits(f);
    const double m = 0.5 * ((double)f + (double)bits_float(fb + 1u));
    return (fb & 1u) ? m : bits_double(double_bits(m) + 1u);
}

struct BoundEntry {
    double c, r;   // C = -ln u0 (u >= 1/2) or ln m0 (u < 1/2), R = 1 / u0 or 1 / m0
};
constexpr int kBoundTabHi = 513;                 // entries for u >= 1/2: i = 0..512
constexpr int kBoundTabN = kBoundTabHi + 1024;   // then u < 1/2: i = 0..1023

// Fills the table (host, glibc log1p).
inline void build_bound_table(BoundEntry* t) {
    for (int i = 0; i < kBoundTabHi; ++i) {
        const double a = (double)i * 0x1p-10;
        t[i].c = -log1p(-a);
        t[i].r = 1.0 / (1.0 - a);
    }
    for (int i = 0; i < 1024; ++i) {
        const double a = (double)i * 0x1p-10;
        t[kBoundTabHi + i].c = log1p(a);
        t[kBoundTabHi + i].r = 1.0 / (1.0 + a);
    }
}

// T = -ln(u) - h/u for u = k 2^-32, k > 0 (h = half an FP64 ulp of u).
// Both branches are evaluated and selected, so a wavefront issues one table
// load per word and no divergent branches.
REWL_AB_HD double accept_threshold(uint32_t k, const BoundEntry* tab) {
    const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
    const bool hi = k >= 0x80000000u;
    // u >= 1/2: u0 = 1 - ia 2^-10, u - u0 = -(j - ia 2^22) 2^-32 (exact)
    const uint32_t j = (uint32_t)(0x100000000ull - (uint64_t)k);
    const uint32_t ia = j >> 22;
    // u < 1/2: k = 2^ex m, m0 = 1 + ib 2^-10, m - m0 = (low 21 bits of mb) 2^-31
    const int lz = __builtin_clz(k | 1u);
    const int ex = 31 - lz;
    const uint32_t mb = k << lz;
    const uint32_t ib = (mb >> 21) & 1023u;
    const BoundEntry e = tab[hi ? ia : kBoundTabHi + ib];
    const double dx = hi ? -((double)(j - (ia << 22)) * 0x1p-32) : (double)(mb & 0x1fffffu) * 0x1p-31;
    const double x = dx * e.r;
    const double n = (double)(32 - ex);
    const double C = hi ? e.c : (n * ln2_hi - e.c) + n * ln2_lo;
    const double hs = hi ? 0x1p-54 : 0x1p-53;
    // ln(1 + x) to x^8
    double p = fma(x, -1.0 / 8.0, 1.0 / 7.0);
    p = fma(x, p, -1.0 / 6.0);
    p = fma(x, p, 1.0 / 5.0);
    p = fma(x, p, -1.0 / 4.0);
    p = fma(x, p, 1.0 / 3.0);
    p = fma(x, p, -1.0 / 2.0);
    p = fma(x, p, 1.0);
    const double T = C - x * p;
    // h / u = hs R / (1 + x) to first order in x
    return T - (hs * e.r) * (1.0 - x);
}

// Largest float f with u < RN(exp(-(double)f)) for u = k 2^-32, k > 0.
REWL_AB_HD float accept_float_max(uint32_t k, const BoundEntry* tab) {
    const double T = accept_threshold(k, tab);
    float f = (float)T;
    if ((double)f >= T) f = bits_float(float_bits(f) - 1u);
    return f;
}

REWL_AB_HD double accept_bound(uint32_t k, const BoundEntry* tab) {
    const float f = accept_float_max(k, tab);
    return float_le_bound(k == 0u ? 0x1.74910cp+9f : f);
}

}  // namespace rewl
