// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), device version.
//
// Ten rounds; each round multiplies counter words 0 and 2 by the Philox
// multipliers (one 32x32->64 product each), and the key is bumped by the
// Weyl constants between rounds. The key of a site update is uniform across
// a wavefront (seed, replica slot), so the compiler keeps the key schedule
// in scalar registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ising {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

// Counter word 2 selects the stream.
constexpr uint32_t kStreamMetropolis = 0u;
constexpr uint32_t kStreamInit = 1u;
constexpr uint32_t kStreamSwap = 2u;

struct u32x4 { uint32_t x, y, z, w; };

// a ^ b ^ c in one instruction (gfx950's three-input bitwise operation,
// truth table 0x96); the compiler emits two XORs for the plain expression.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += kPhiloxW0; k1 += kPhiloxW1; }
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        // rounds 0 and 1: with counter words 1..3 wavefront-uniform (sweep,
        // stream, seed) one operand of each XOR besides the key is uniform or
        // zero, and the plain form lets them fold into scalar operations (the
        // three-input form would need a vector copy of one of two scalars)
        if (r < 2) c = u32x4{ hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0 };
        else c = u32x4{ xor3(hi1, c.y, k0), lo1, xor3(hi0, c.w, k1), lo0 };
    }
    return c;
}

// 24-bit uniforms, exactly representable in FP32 (and identical in FP64).
__device__ __forceinline__ float u24f(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }        // [0,1)
__device__ __forceinline__ float s24f(uint32_t x) { return (float)(x >> 8) * 0x1p-23f - 1.0f; } // [-1,1)

} // namespace ising
