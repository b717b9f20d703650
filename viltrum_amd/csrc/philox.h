/*
 * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the
 * published multipliers and Weyl key increments.  Stateless: every residual sample
 * draws from counter (sample, bin, integral, block) under key (seed), so results do not
 * depend on launch shape, thread count or sharding.
 */
#ifndef VMX_PHILOX_H
#define VMX_PHILOX_H

#include <stdint.h>

#include "../../include/viltrum_amd/integrands.h"

namespace vmx {

struct U4 {
    uint32_t x, y, z, w;
};

VMX_HD inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += W0;
            k1 += W1;
        }
        // one 32x32->64 product per multiplier: a single v_mad_u64_u32 on CDNA instead of
        // a v_mul_hi_u32 + v_mul_lo_u32 pair (same bits, ~25% faster per call)
        const uint64_t p0 = uint64_t(M0) * uint64_t(c.x);
        const uint64_t p1 = uint64_t(M1) * uint64_t(c.z);
        const U4 n = {uint32_t(p1 >> 32) ^ c.y ^ k0, uint32_t(p1), uint32_t(p0 >> 32) ^ c.w ^ k1, uint32_t(p0)};
        c = n;
    }
    return c;
}

/* [0, 1) from 24 random bits */
VMX_HD inline float u01f(uint32_t u) { return float(u >> 8) * (1.0f / 16777216.0f); }
/* [0, 1) from 53 random bits */
VMX_HD inline double u01d(uint32_t a, uint32_t b) {
    return double(((uint64_t(a) << 32) | b) >> 11) * (1.0 / 9007199254740992.0);
}
/* uniform integer in [0, n) by 64-bit multiply-shift of 32 random bits: bias < n / 2^32 */
VMX_HD inline uint32_t uniform_index(uint32_t u, uint32_t n) { return uint32_t((uint64_t(u) * n) >> 32); }

}  // namespace vmx

#endif
