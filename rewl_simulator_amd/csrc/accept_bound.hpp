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
// within 0.002 ulp of the rounding midpoint u + h, and T carries the error of
// one FP64 log, which moves f only when a float lies within an ulp of T.
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

// Largest float f with u < RN(exp(-(double)f)) for u = k 2^-32, k > 0.
REWL_AB_HD float accept_float_max(uint32_t k) {
    const double u = (double)k * 0x1p-32;
    int e;
    (void)frexp(u, &e);                      // u = m 2^e, m in [0.5, 1): ulp(u) = 2^(e-53)
    const double h = ldexp(1.0, e - 54);
    const double T = -log(u) - h / u;
    float f = (float)T;
    if ((double)f >= T) f = bits_float(float_bits(f) - 1u);
    return f;
}

REWL_AB_HD double accept_bound(uint32_t k) {
    return float_le_bound(k == 0u ? 0x1.74910cp+9f : accept_float_max(k));
}

}  // namespace rewl
