// es_calib.hip -- peak calibration kernels (C ABI of include/es_calib.h): INT32 VALU adds and
// LDS 8-byte reads.
//
// Each thread runs 8 independent chains a_j += b (v_add_u32 written in inline asm so the
// compiler cannot fold the chain into a multiply); 8 chains keep each SIMD issuing one add
// per cycle despite the VALU dependency latency.  ops = threads x iters x 256.
#include "es_common.h"
#include "es_calib.h"

namespace {

__global__ void __launch_bounds__(256) es_valu_chain_kernel(uint32_t iters, uint32_t seed,
                                                            uint32_t* __restrict__ sink) {
  uint32_t a0 = threadIdx.x, a1 = a0 ^ 1u, a2 = a0 ^ 2u, a3 = a0 ^ 3u;
  uint32_t a4 = a0 ^ 4u, a5 = a0 ^ 5u, a6 = a0 ^ 6u, a7 = a0 ^ 7u;
  const uint32_t b = seed | 1u;
  for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 32; ++r) {
      asm volatile(
          "v_add_u32 %0, %0, %8\n\t"
          "v_add_u32 %1, %1, %8\n\t"
          "v_add_u32 %2, %2, %8\n\t"
          "v_add_u32 %3, %3, %8\n\t"
          "v_add_u32 %4, %4, %8\n\t"
          "v_add_u32 %5, %5, %8\n\t"
          "v_add_u32 %6, %6, %8\n\t"
          "v_add_u32 %7, %7, %8"
          : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
          : "v"(b));
    }
  }
  const uint32_t x = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
  if (x == 0x9e3779b9u) sink[blockIdx.x] = x;  // keeps the chains live
}

// LDS read rate: every lane of a wave reads an 8-byte row (ds_read_b64) of a 32 KiB table in
// LDS, eight reads per round.  The row is the lane's offset plus a wave-uniform step, the
// access pattern of the sieve's mask reads: mode 0 gives the 64 lanes of a wave 64
// consecutive rows (each half-wave covers the 64 banks once: no conflicts), mode 1 a fixed
// pseudo-random row per lane (the sieve's random per-lane mask rows).
constexpr uint32_t kLdsRows = 4096;  // 8-byte rows: 32 KiB

__global__ void __launch_bounds__(1024) es_lds_read_kernel(uint32_t iters, uint32_t mode,
                                                           uint32_t* __restrict__ sink) {
  __shared__ uint2 tab[kLdsRows];
  for (uint32_t i = threadIdx.x; i < kLdsRows; i += blockDim.x)
    tab[i] = make_uint2(i * 2654435761u, i ^ 0x5bd1e995u);
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t off = wave * 64u + lane;
  if (mode == 1) {
    uint32_t h = (threadIdx.x + 1u) * 0x9e3779b1u ^ (blockIdx.x * 0x85ebca6bu);
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    off = h;
  }
  // LDS byte address of the lane's row in the lower half of the table; the eight reads of a
  // round sit 2 KiB apart (immediate offsets), and the round's uniform step moves whole groups
  // of 64 rows, so every read of a wave keeps the lane pattern's banks.  The reads are written
  // out as ds_read_b64 (the compiler would pair them into ds_read2st64_b64, another
  // instruction than the sieve's); the block waits for them before it ends, and its outputs
  // are early-clobber, so no read can land in the address register while others issue.
  const uint32_t lane_b = (off & (kLdsRows / 2 - 1u)) * 8u;
  const uint32_t tab_b = (uint32_t)(uintptr_t)tab;
  uint32_t acc = 0;
  uint32_t step = 0;
  for (uint32_t it = 0; it < iters; ++it) {
    const uint32_t a = tab_b + ((lane_b + step) & (kLdsRows * 4u - 1u));
    uint64_t v0, v1, v2, v3, v4, v5, v6, v7;
    asm volatile(
        "ds_read_b64 %0, %8\n\t"
        "ds_read_b64 %1, %8 offset:2048\n\t"
        "ds_read_b64 %2, %8 offset:4096\n\t"
        "ds_read_b64 %3, %8 offset:6144\n\t"
        "ds_read_b64 %4, %8 offset:8192\n\t"
        "ds_read_b64 %5, %8 offset:10240\n\t"
        "ds_read_b64 %6, %8 offset:12288\n\t"
        "ds_read_b64 %7, %8 offset:14336\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3), "=&v"(v4), "=&v"(v5), "=&v"(v6),
          "=&v"(v7)
        : "v"(a)
        : "memory");
    acc = (acc ^ (uint32_t)v0) + (uint32_t)(v0 >> 32);
    acc = (acc ^ (uint32_t)v1) + (uint32_t)(v1 >> 32);
    acc = (acc ^ (uint32_t)v2) + (uint32_t)(v2 >> 32);
    acc = (acc ^ (uint32_t)v3) + (uint32_t)(v3 >> 32);
    acc = (acc ^ (uint32_t)v4) + (uint32_t)(v4 >> 32);
    acc = (acc ^ (uint32_t)v5) + (uint32_t)(v5 >> 32);
    acc = (acc ^ (uint32_t)v6) + (uint32_t)(v6 >> 32);
    acc = (acc ^ (uint32_t)v7) + (uint32_t)(v7 >> 32);
    step += 512u * 5u;
  }
  if (acc == 0x9e3779b9u) sink[blockIdx.x] = acc;  // keeps the reads live
}

// HBM streaming read: each thread reads 16-byte vectors (global_load_dwordx4) at
// grid-stride positions, U of them in flight before it folds them into its checksum; NT marks
// the loads nontemporal (streamed past the caches).  A launch reads every byte once.
typedef uint32_t es_v4u __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ void __launch_bounds__(512) es_hbm_read_kernel(const es_v4u* __restrict__ src,
                                                          uint64_t n16,
                                                          uint32_t* __restrict__ sink) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  es_v4u acc = {0u, 0u, 0u, 0u};
  for (; i + (uint64_t)(U - 1) * stride < n16; i += (uint64_t)U * stride) {
    es_v4u v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = NT ? __builtin_nontemporal_load(src + i + (uint64_t)u * stride)
                : src[i + (uint64_t)u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) acc ^= v[u];
  }
  for (; i < n16; i += stride) acc ^= src[i];
  const uint32_t x = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (x == 0x9e3779b9u) sink[threadIdx.x] = x;  // keeps the loads live
}

}  // namespace

namespace {

int run_lds(uint32_t blocks, uint32_t iters, uint32_t mode, uint32_t* sink, hipEvent_t e0,
            hipEvent_t e1, double* ms) {
  hipLaunchKernelGGL(es_lds_read_kernel, dim3(blocks), dim3(1024), 0, nullptr, 1u, mode, sink);
  ES_HIP(hipGetLastError());  // warm-up launch
  ES_HIP(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(es_lds_read_kernel, dim3(blocks), dim3(1024), 0, nullptr, iters, mode, sink);
  ES_HIP(hipGetLastError());
  ES_HIP(hipEventRecord(e1, nullptr));
  ES_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  ES_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t;
  return ES_OK;
}

int run_valu(uint32_t blocks, uint32_t iters, uint32_t* sink, hipEvent_t e0, hipEvent_t e1,
             double* ms) {
  hipLaunchKernelGGL(es_valu_chain_kernel, dim3(blocks), dim3(256), 0, nullptr, 1u, 7u, sink);
  ES_HIP(hipGetLastError());  // warm-up launch
  ES_HIP(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(es_valu_chain_kernel, dim3(blocks), dim3(256), 0, nullptr, iters, 7u, sink);
  ES_HIP(hipGetLastError());
  ES_HIP(hipEventRecord(e1, nullptr));
  ES_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  ES_HIP(hipEventElapsedTime(&t, e0, e1));
  *ms = t;
  return ES_OK;
}

}  // namespace

extern "C" int es_calib_valu(uint32_t blocks, uint32_t iters, double* ms, uint64_t* ops) {
  if (!ms || !ops || blocks == 0 || iters == 0) return es::fail(ES_ERR_INVALID, "bad argument");
  uint32_t* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = ES_OK;
  if (hipMalloc(&sink, (size_t)blocks * 4) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = es::fail(ES_ERR_DEVICE, "es_calib_valu: allocation failed");
  if (rc == ES_OK) rc = run_valu(blocks, iters, sink, e0, e1, ms);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (sink) (void)hipFree(sink);
  if (rc == ES_OK) *ops = (uint64_t)blocks * 256u * iters * 256u;
  return rc;
}

extern "C" int es_calib_lds(uint32_t blocks, uint32_t iters, uint32_t mode, double* ms,
                            uint64_t* bytes) {
  if (!ms || !bytes || blocks == 0 || iters == 0 || mode > 1)
    return es::fail(ES_ERR_INVALID, "bad argument");
  uint32_t* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = ES_OK;
  if (hipMalloc(&sink, (size_t)blocks * 4) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = es::fail(ES_ERR_DEVICE, "es_calib_lds: allocation failed");
  if (rc == ES_OK) rc = run_lds(blocks, iters, mode, sink, e0, e1, ms);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (sink) (void)hipFree(sink);
  if (rc == ES_OK) *bytes = (uint64_t)blocks * 1024u * iters * 8u * 8u;
  return rc;
}

extern "C" int es_calib_hbm_read(const void* src, uint64_t bytes, uint32_t blocks,
                                 uint32_t threads, uint32_t unroll, uint32_t nontemporal,
                                 uint32_t reps, double* ms) {
  if (!src || !ms || bytes < 16 || bytes % 16 || blocks == 0 || reps == 0 ||
      (threads != 256 && threads != 512) || ((uintptr_t)src & 15u))
    return es::fail(ES_ERR_INVALID, "es_calib_hbm_read: bad argument");
  using Fn = void (*)(const es_v4u*, uint64_t, uint32_t*);
  Fn fn = nullptr;
  if (unroll == 1) fn = nontemporal ? es_hbm_read_kernel<1, true> : es_hbm_read_kernel<1, false>;
  if (unroll == 4) fn = nontemporal ? es_hbm_read_kernel<4, true> : es_hbm_read_kernel<4, false>;
  if (unroll == 8) fn = nontemporal ? es_hbm_read_kernel<8, true> : es_hbm_read_kernel<8, false>;
  if (!fn) return es::fail(ES_ERR_INVALID, "es_calib_hbm_read: unroll must be 1, 4 or 8");
  uint32_t* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = ES_OK;
  if (hipMalloc(&sink, 512 * 4) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = es::fail(ES_ERR_DEVICE, "es_calib_hbm_read: allocation failed");
  const es_v4u* p = reinterpret_cast<const es_v4u*>(src);
  const uint64_t n16 = bytes / 16;
  if (rc == ES_OK) {
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(threads), 0, nullptr, p, n16, sink);  // warm-up
    if (hipGetLastError() != hipSuccess) rc = es::fail(ES_ERR_DEVICE, "es_calib_hbm_read: launch");
  }
  double best = 1e30;
  for (uint32_t r = 0; rc == ES_OK && r < reps; ++r) {
    float t = 0.f;
    if (hipEventRecord(e0, nullptr) != hipSuccess) rc = es::fail(ES_ERR_DEVICE, "event");
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(threads), 0, nullptr, p, n16, sink);
    if (rc == ES_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&t, e0, e1) != hipSuccess))
      rc = es::fail(ES_ERR_DEVICE, "es_calib_hbm_read: launch");
    best = std::min(best, (double)t);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (sink) (void)hipFree(sink);
  if (rc == ES_OK) *ms = best;
  return rc;
}
