// Philox4x32-10 counter RNG on the device (Salmon, Moraes, Dror, Shaw,
// SC'11): draw k of stream t under `seed` is a pure function of (seed, t, k);
// key = seed, counter = (k_lo, k_hi, t_lo, t_hi); the uniform double in
// [0, 1) is built from the top 53 bits of the first two output words.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace yq {

__device__ __forceinline__ void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t o0 = hi1 ^ c[1] ^ k[0], o2 = hi0 ^ c[3] ^ k[1];
    c[0] = o0;
    c[1] = lo1;
    c[2] = o2;
    c[3] = lo0;
}

__device__ __forceinline__ double philox_uniform_dev(uint64_t seed, uint64_t traj, uint64_t kk) {
    uint32_t c[4] = {(uint32_t)kk, (uint32_t)(kk >> 32), (uint32_t)traj, (uint32_t)(traj >> 32)};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    const uint64_t bits = ((uint64_t)(c[0] >> 5) << 26) | (uint64_t)(c[1] >> 6);
    return (double)bits * (1.0 / 9007199254740992.0);
}

}  // namespace yq
