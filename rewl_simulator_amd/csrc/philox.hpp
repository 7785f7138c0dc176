// Counter-based random numbers for the walkers.
//
// Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  The reference draws
// from one std::mt19937_64 per MPI rank (random_number_generators.hpp:14-59);
// a counter-based generator lets every lane derive its uniforms from
// (seed, walker, position) with no state to store or exchange.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define REWL_HD __host__ __device__ __forceinline__
#else
#define REWL_HD inline
#endif

namespace rewl {

struct u32x4 { uint32_t x, y, z, w; };

REWL_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x;
        const uint64_t p1 = (uint64_t)M1 * c.z;
        u32x4 o;
        o.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        o.y = (uint32_t)p1;
        o.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        o.w = (uint32_t)p0;
        c = o;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Stream domains of a walker's Philox counter space.
enum : uint32_t { DOMAIN_SWEEP = 0, DOMAIN_INIT = 1, DOMAIN_EXCH = 2 };

// libstdc++'s generate_canonical<double, 53> over a 64-bit engine: one engine
// call, double(x) * 2^-64, moved below 1 when the conversion rounds up to 2^64.
REWL_HD double canonical_from_u64(uint64_t x) {
    double u = (double)x * 0x1p-64;
    if (u >= 1.0) u = 0x1.fffffffffffffp-1;
    return u;
}

REWL_HD uint64_t philox_u64(uint32_t k0, uint32_t k1, uint32_t domain, uint32_t walker, uint64_t index) {
    const u32x4 o = philox4x32_10(u32x4{(uint32_t)index, (uint32_t)(index >> 32), domain, walker}, k0, k1);
    return (uint64_t)o.x | ((uint64_t)o.y << 32);
}

}  // namespace rewl
