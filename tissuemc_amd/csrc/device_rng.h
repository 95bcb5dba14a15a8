// Philox4x32-10 for gfx950, the counter-based generator behind every photon's random stream.
//
// The stream layout is the reference's (include/tissuemc/philox.hpp:59-83): counter words 0-1
// hold the 64-bit photon id, word 2 counts the 128-bit blocks drawn so far, word 3 is a stream
// word (0 in the reference), and the key is the 64-bit run seed. Words are consumed in order
// v[0], v[1], v[2], v[3], then the next block. Round structure and constants follow Salmon et
// al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11), the same published algorithm the
// reference implements at philox.hpp:26-57.
//
// On gfx950 each round is two v_mad_u64_u32 (32x32->64 products) and two v_bitop3_b32
// (three-input XORs); the key schedule is wave-uniform, so it stays in scalar registers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmc {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

struct Block4 {
    uint32_t v0, v1, v2, v3;
};

// a ^ b ^ c. On the device this is one v_bitop3_b32 (truth table 0x96), a gfx950 instruction
// the compiler does not form from two XORs by itself.
__host__ __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

__host__ __device__ __forceinline__ Block4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                         uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t p0 = static_cast<uint64_t>(kPhiloxM0) * c0;
        const uint64_t p1 = static_cast<uint64_t>(kPhiloxM1) * c2;
        const uint32_t n0 = xor3(static_cast<uint32_t>(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3(static_cast<uint32_t>(p0 >> 32), c3, k1);
        c1 = static_cast<uint32_t>(p1);
        c3 = static_cast<uint32_t>(p0);
        c0 = n0;
        c2 = n2;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return Block4{c0, c1, c2, c3};
}

// Uniform deviate in (0, 1) from one 32-bit word, keeping 23 bits: ((x >> 9) + 0.5) * 2^-23,
// the reference's float mapping (include/tissuemc/philox.hpp:133-136), bit for bit. The top 23
// bits become the mantissa of a float f = 1 + (x >> 9) 2^-23 in [1, 2) (one v_alignbit_b32
// shifting the exponent bits 0x7F in above them), and f - (1 - 2^-24) is exact (both operands
// within a factor of two of each other): two instructions instead of a shift, a conversion and
// an fma. `expo` is kUnitExponent; a caller in a loop passes it in a vector register (the
// instruction cannot take it as a literal), which keeps it out of the scalar registers.
constexpr uint32_t kUnitExponent = 0x7Fu;
constexpr float kUnitOffset = 0x1.fffffep-1f;  // 1 - 2^-24
__device__ __forceinline__ float unit_open(uint32_t x, uint32_t expo = kUnitExponent) {
    return __uint_as_float(__builtin_amdgcn_alignbit(expo, x, 9u)) - kUnitOffset;
}

// unit_open(x) < 0.1f exactly when x < kUnitBelowTenth: (m + 0.5) 2^-23 < 0.1f holds for the
// 23-bit m = x >> 9 up to 838860, and 838861 * 512 = 429496832.
constexpr uint32_t kUnitBelowTenth = 429496832u;

}  // namespace tmc
