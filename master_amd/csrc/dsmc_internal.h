// dsmc_internal.h -- host-side helpers shared by the library's translation units.
#pragma once
#include <stdint.h>

namespace dsmc {

// Philox4x32-10 on the host (same counter/key layout as the device).
inline void host_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                        uint32_t out[4]) {
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

inline double host_u01(uint32_t x) { return (double)((x >> 9) * 2u + 1u) * (1.0 / 16777216.0); }

}  // namespace dsmc
