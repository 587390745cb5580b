// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel Random Numbers: As Easy as 1, 2, 3", SC'11), usable from host
// and device code.  Replaces the reference's cuRAND XORWOW state seeded by
// thread id (non-stream/main.cu:40-41, 109): every draw here is a pure
// function of (key, logical coordinates), so results do not depend on the
// launch geometry or on which kernel variant runs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace csaw {

// Counter domain tags, bits 28..31 of counter word 3.
enum : uint32_t {
  kDomSeeds = 1,
  kDomRmat = 2,
  kDomWeights = 3,
  kDomSelect = 4,
  kDomWalk = 5,
  kDomMdrw = 6,
  kDomMix = 7,
  kDomExplore = 8
};

struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                     uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return U4{c0, c1, c2, c3};
}

// 64-bit target t = (o1 << 32) | o0.
__host__ __device__ __forceinline__ uint64_t philox_t64(const U4& o) {
  return ((uint64_t)o.y << 32) | o.x;
}

// floor(a * b / 2^64), exact.
__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// floor(a * b / 2^64) for a 32-bit b, exact with two 32x32->64 products:
// a * b = (a_hi * b) * 2^32 + a_lo * b, and a_hi * b + floor(a_lo * b / 2^32)
// < 2^64, so the floor of the whole divides out in two steps.
__host__ __device__ __forceinline__ uint64_t mulhi64_32(uint64_t a, uint32_t b) {
  const uint64_t lo = (uint64_t)(uint32_t)a * b;
  const uint64_t hi = (a >> 32) * b;
  return (hi + (lo >> 32)) >> 32;
}

// Uniform float in [0,1) with 24 random bits: (o0 >> 8) * 2^-24, exact.
__host__ __device__ __forceinline__ float philox_u24(const U4& o) {
  return (float)(o.x >> 8) * 5.9604644775390625e-08f;
}

}  // namespace csaw
