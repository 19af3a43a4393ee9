// es_calib.hip -- INT32 VALU peak calibration kernel (C ABI of include/es_calib.h).
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

}  // namespace

namespace {

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
