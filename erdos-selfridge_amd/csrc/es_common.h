// es_common.h -- helpers shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/es_sieve.h"

typedef unsigned __int128 u128;

// Loads through this pointer type are from the constant address space: the compiler may use
// scalar (SMEM) loads when the address is wave-uniform, and assumes the data is read-only for
// the kernel's lifetime.
typedef __attribute__((address_space(4))) const uint32_t cu32_t;

namespace es {

std::string& last_error();
int fail(int code, const std::string& msg);

__host__ __device__ inline u128 mk128(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }
inline uint32_t mod_u128(u128 x, uint32_t q) { return (uint32_t)(x % q); }

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
