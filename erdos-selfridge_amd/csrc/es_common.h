// es_common.h -- helpers shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/es_sieve.h"

typedef unsigned __int128 u128;

// Loads through this pointer type are from the constant address space: the compiler may use
// scalar (SMEM) loads when the address is wave-uniform, and assumes the data is read-only for
// the kernel's lifetime.
typedef __attribute__((address_space(4))) const uint32_t cu32_t;
typedef __attribute__((address_space(4))) const double cf64_t;

namespace es {

std::string& last_error();
int fail(int code, const std::string& msg);

__host__ __device__ inline u128 mk128(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }
inline uint32_t mod_u128(u128 x, uint32_t q) { return (uint32_t)(x % q); }
// a^-1 mod m (extended Euclid), or 0 when gcd(a, m) != 1 or m < 2
inline uint32_t inv_mod(uint32_t a, uint32_t m) {
  if (m < 2) return 0;
  int64_t r0 = m, r1 = a % m, s0 = 0, s1 = 1;
  while (r1) {
    const int64_t qt = r0 / r1, r2 = r0 - qt * r1, s2 = s0 - qt * s1;
    r0 = r1, r1 = r2, s0 = s1, s1 = s2;
  }
  if (r0 != 1) return 0;
  return (uint32_t)(s0 < 0 ? s0 + m : s0);
}

}  // namespace es

#define ES_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t err_ = (call);                                                                 \
    if (err_ != hipSuccess)                                                                   \
      return ::es::fail(ES_ERR_DEVICE, std::string(#call) + " failed: " + hipGetErrorString(err_)); \
  } while (0)

__device__ __forceinline__ bool es_ge128(uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) {
  return (ah > bh) || (ah == bh && al >= bl);
}

// x mod q of a 128-bit x from 2^32, 2^64, 2^96 mod q (q < 2^24 keeps the sum below 2^58).
__device__ __forceinline__ uint32_t es_mod128(uint64_t lo, uint64_t hi, uint32_t q, uint32_t p1,
                                              uint32_t p2, uint32_t p3) {
  const uint64_t x = (uint64_t)(uint32_t)lo + (uint64_t)(uint32_t)(lo >> 32) * p1 +
                     (uint64_t)(uint32_t)hi * p2 + (uint64_t)(uint32_t)(hi >> 32) * p3;
  return (uint32_t)(x % q);
}
