// kg_philox.h -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable from
// host and device code so that every backend draws the same noise.
//
// Noise of cell (i, j) at step s of ensemble member `inst`:
//   counter = (i >> 2, j, s, inst), key = (seed lo, seed hi),
//   chi = (word[i & 3] >> 8) * 2^-24 - 1/2  in [-1/2, 1/2), exact in float.
// Four neighbouring cells of a row share one Philox block.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kg {

struct u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ void philox_mulhilo(uint32_t a, uint32_t b, uint32_t &hi,
                                                        uint32_t &lo)
{
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        philox_mulhilo(M0, c.x, hi0, lo0);
        philox_mulhilo(M1, c.z, hi1, lo1);
        c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

__host__ __device__ __forceinline__ uint32_t select_word(const u32x4 &v, uint32_t k)
{
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}

// chi in [-1/2, 1/2) from a 32-bit word; exact in float and double.
template <typename R>
__host__ __device__ __forceinline__ R noise_from_word(uint32_t w)
{
    return (R)(w >> 8) * (R)(1.0 / 16777216.0) - (R)0.5;
}

}  // namespace kg
