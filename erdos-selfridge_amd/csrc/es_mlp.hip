// es_mlp.hip -- the wheel-planner surrogate: a 64-128-128-2 FP32 multilayer perceptron with
// ReLU hidden layers, trained with Adam on mean squared error, all on the GPU.
//
// No reference counterpart: the reference plans wheels with an analytic knapsack only
// (sieve/gpu_sieve_v9.py:146-227).  The surrogate predicts, per candidate wheel, the log
// survivor density and the measured ns per residue position, so that many candidate wheels
// can be ranked cheaply before a short GPU probe picks among the best few.
//
// Every layer is a small GEMM run on the matrix cores: es_gemm_kernel computes one 16x16
// output tile per wave with v_mfma_f32_16x16x4f32 (lane l supplies A[l % 16][k0 + l / 16] and
// B[k0 + l / 16][l % 16]; accumulator element j of lane l is D[4 (l / 16) + j][l % 16]).
// Operands are addressed through row/column strides, so transposed products (the backward
// pass) need no copies.  Epilogue flags add a bias, apply ReLU, or multiply by the ReLU mask
// of a stored activation.  Batches are padded to a multiple of 16 rows with zero loss
// gradient, and the 2-wide output layer is held 16 wide with zero columns that stay zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "es_common.h"
#include "../../include/es_mlp.h"

using es::fail;

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kIn = ES_MLP_IN, kH = ES_MLP_HIDDEN, kOut = ES_MLP_OUT, kOutPad = 16;
// packed parameter layout (floats): W1 [kIn x kH], b1 [kH], W2 [kH x kH], b2 [kH],
// W3 [kH x kOutPad], b3 [kOutPad]; weights row-major as (input, output)
constexpr uint32_t oW1 = 0, ob1 = oW1 + kIn * kH, oW2 = ob1 + kH, ob2 = oW2 + kH * kH,
                   oW3 = ob2 + kH, ob3 = oW3 + kH * kOutPad, kParams = ob3 + kOutPad;
static_assert(kParams == ES_MLP_PARAMS, "packed parameter count");

enum : uint32_t { kBias = 1u, kRelu = 2u, kMask = 4u };

struct GemmArgs {
  const float* A;  // A(i, k) = A[i * a_rs + k * a_cs]
  const float* B;  // B(k, j) = B[k * b_rs + j * b_cs]
  float* C;        // C(i, j) = C[i * ldc + j]
  const float* bias;  // [N] when kBias
  const float* mask;  // mask(i, j) = mask[i * ldc + j] > 0 keeps the value, when kMask
  uint32_t M, N, K, a_rs, a_cs, b_rs, b_cs, ldc, flags;
};

// One wave per 16x16 tile of C, four waves per workgroup.  M, N multiples of 16, K of 4.
__global__ void __launch_bounds__(256) es_gemm_kernel(GemmArgs g) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
  const uint32_t tiles_n = g.N / 16u;
  if (tile >= (g.M / 16u) * tiles_n) return;
  const uint32_t i0 = (tile / tiles_n) * 16u, j0 = (tile % tiles_n) * 16u;
  const uint32_t r = lane & 15u, kk = lane >> 4;
  const float* a = g.A + (size_t)(i0 + r) * g.a_rs + (size_t)kk * g.a_cs;
  const float* b = g.B + (size_t)kk * g.b_rs + (size_t)(j0 + r) * g.b_cs;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (uint32_t k0 = 0; k0 < g.K; k0 += 4u) {
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[(size_t)k0 * g.a_cs], b[(size_t)k0 * g.b_rs],
                                                acc, 0, 0, 0);
  }
  const uint32_t j = j0 + r;
  const float bj = (g.flags & kBias) ? g.bias[j] : 0.f;
#pragma unroll
  for (uint32_t e = 0; e < 4u; ++e) {
    const uint32_t i = i0 + 4u * kk + e;
    float v = acc[e] + bj;
    if (g.flags & kRelu) v = v > 0.f ? v : 0.f;
    if (g.flags & kMask) v = g.mask[(size_t)i * g.ldc + j] > 0.f ? v : 0.f;
    g.C[(size_t)i * g.ldc + j] = v;
  }
}

// dY = 2 w_o (Y - T) / n on rows < n (zero on padding rows and padded outputs); loss terms
// per row into part[row].
__global__ void es_loss_grad_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                    float* __restrict__ dY, float* __restrict__ part, uint32_t n,
                                    uint32_t rows, float w0, float w1) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  float l = 0.f;
  for (uint32_t o = 0; o < kOutPad; ++o) {
    float g = 0.f;
    if (i < n && o < kOut) {
      const float w = o == 0 ? w0 : w1;
      const float d = Y[i * kOutPad + o] - T[i * kOut + o];
      l += w * d * d;
      g = 2.f * w * d / (float)n;
    }
    dY[i * kOutPad + o] = g;
  }
  part[i] = i < n ? l : 0.f;
}

// out[j] = sum_i in[i * cols + j], one thread per column (rows is a few thousand at most).
__global__ void es_colsum_kernel(const float* __restrict__ in, float* __restrict__ out,
                                 uint32_t rows, uint32_t cols) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (uint32_t i = 0; i < rows; ++i) s += in[(size_t)i * cols + j];
  out[j] = s;
}

// Adam with bias correction (m_hat = m / (1 - b1^t), v_hat = v / (1 - b2^t)).
__global__ void es_adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                               float* __restrict__ m, float* __restrict__ v, uint32_t n, float lr,
                               float b1, float b2, float eps, float c1, float c2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Gather rows idx[0..n) of the resident dataset into the batch buffers (X 64 wide, T 2 wide);
// rows n..rows are zero.
__global__ void es_gather_kernel(const float* __restrict__ DX, const float* __restrict__ DT,
                                 const uint32_t* __restrict__ idx, float* __restrict__ X,
                                 float* __restrict__ T, uint32_t n, uint32_t rows) {
  const uint32_t i = blockIdx.x;
  if (i >= rows) return;
  const bool live = i < n;
  const uint32_t s = live ? idx[i] : 0u;
  for (uint32_t c = threadIdx.x; c < kIn; c += blockDim.x)
    X[(size_t)i * kIn + c] = live ? DX[(size_t)s * kIn + c] : 0.f;
  if (threadIdx.x < kOut) T[(size_t)i * kOut + threadIdx.x] = live ? DT[(size_t)s * kOut + threadIdx.x] : 0.f;
}

}  // namespace

struct es_mlp {
  uint32_t cap = 0;  // batch rows (multiple of 16)
  hipStream_t stream = nullptr;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  float *X = nullptr, *T = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr,
        *dH1 = nullptr, *dH2 = nullptr, *part = nullptr;
  float *DX = nullptr, *DT = nullptr;  // resident training set
  uint32_t *idx = nullptr;
  uint64_t n_data = 0;
  uint64_t step = 0;
  std::vector<void*> allocs;
  float* alloc(size_t floats) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(floats, 16) * 4) != hipSuccess) return nullptr;
    allocs.push_back(p);
    return (float*)p;
  }
  ~es_mlp() {
    if (stream) hipStreamSynchronize(stream);
    for (void* p : allocs) hipFree(p);
    hipFree(DX);
    hipFree(DT);
    hipFree(idx);
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

int gemm(es_mlp* n, const float* A, uint32_t a_rs, uint32_t a_cs, const float* B, uint32_t b_rs,
         uint32_t b_cs, float* C, uint32_t ldc, uint32_t M, uint32_t N, uint32_t K,
         uint32_t flags, const float* bias = nullptr, const float* mask = nullptr) {
  if (M % 16 || N % 16 || K % 4 || ldc < N) return fail(ES_ERR_INVALID, "gemm shape");
  GemmArgs g{A, B, C, bias, mask, M, N, K, a_rs, a_cs, b_rs, b_cs, ldc, flags};
  const uint32_t tiles = (M / 16) * (N / 16);
  hipLaunchKernelGGL(es_gemm_kernel, dim3((tiles + 3) / 4), dim3(256), 0, n->stream, g);
  ES_HIP(hipGetLastError());
  return ES_OK;
}

int forward_rows(es_mlp* n, uint32_t rows) {
  const float* P = n->params;
  int rc;
  if ((rc = gemm(n, n->X, kIn, 1, P + oW1, kH, 1, n->H1, kH, rows, kH, kIn, kBias | kRelu,
                 P + ob1)))
    return rc;
  if ((rc = gemm(n, n->H1, kH, 1, P + oW2, kH, 1, n->H2, kH, rows, kH, kH, kBias | kRelu,
                 P + ob2)))
    return rc;
  return gemm(n, n->H2, kH, 1, P + oW3, kOutPad, 1, n->Y, kOutPad, rows, kOutPad, kH, kBias,
              P + ob3);
}

// Gradients of the loss over the batch already in X/T (n real rows of `rows`) into n->grads;
// the loss into *loss when non-null (synchronises).
int backward_rows(es_mlp* n, uint32_t nreal, uint32_t rows, float w0, float w1, float* loss) {
  int rc;
  if ((rc = forward_rows(n, rows))) return rc;
  hipLaunchKernelGGL(es_loss_grad_kernel, dim3((rows + 255) / 256), dim3(256), 0, n->stream,
                     n->Y, n->T, n->dY, n->part, nreal, rows, w0, w1);
  ES_HIP(hipGetLastError());
  float* G = n->grads;
  const float* P = n->params;
  // output layer
  if ((rc = gemm(n, n->H2, 1, kH, n->dY, kOutPad, 1, G + oW3, kOutPad, kH, kOutPad, rows, 0)))
    return rc;
  hipLaunchKernelGGL(es_colsum_kernel, dim3(1), dim3(64), 0, n->stream, n->dY, G + ob3, rows,
                     kOutPad);
  if ((rc = gemm(n, n->dY, kOutPad, 1, P + oW3, 1, kOutPad, n->dH2, kH, rows, kH, kOutPad, kMask,
                 nullptr, n->H2)))
    return rc;
  // hidden layer 2
  if ((rc = gemm(n, n->H1, 1, kH, n->dH2, kH, 1, G + oW2, kH, kH, kH, rows, 0))) return rc;
  hipLaunchKernelGGL(es_colsum_kernel, dim3(1), dim3(kH), 0, n->stream, n->dH2, G + ob2, rows, kH);
  if ((rc = gemm(n, n->dH2, kH, 1, P + oW2, 1, kH, n->dH1, kH, rows, kH, kH, kMask, nullptr,
                 n->H1)))
    return rc;
  // hidden layer 1
  if ((rc = gemm(n, n->X, 1, kIn, n->dH1, kH, 1, G + oW1, kH, kIn, kH, rows, 0))) return rc;
  hipLaunchKernelGGL(es_colsum_kernel, dim3(1), dim3(kH), 0, n->stream, n->dH1, G + ob1, rows, kH);
  ES_HIP(hipGetLastError());
  if (loss) {
    std::vector<float> part(rows);
    ES_HIP(hipMemcpyAsync(part.data(), n->part, rows * 4, hipMemcpyDeviceToHost, n->stream));
    ES_HIP(hipStreamSynchronize(n->stream));
    double s = 0;
    for (uint32_t i = 0; i < nreal; ++i) s += part[i];
    *loss = (float)(s / std::max<uint32_t>(nreal, 1));
  }
  return ES_OK;
}

int adam(es_mlp* n, float lr, float b1, float b2, float eps) {
  n->step += 1;
  const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
  hipLaunchKernelGGL(es_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, n->stream,
                     n->params, n->grads, n->m, n->v, kParams, lr, b1, b2, eps, c1, c2);
  ES_HIP(hipGetLastError());
  return ES_OK;
}

uint32_t pad16(uint32_t x) { return (x + 15u) & ~15u; }

int upload_batch(es_mlp* n, const float* X, const float* T, uint32_t cnt, uint32_t rows) {
  std::vector<float> x((size_t)rows * kIn, 0.f), t((size_t)rows * kOut, 0.f);
  std::memcpy(x.data(), X, (size_t)cnt * kIn * 4);
  if (T) std::memcpy(t.data(), T, (size_t)cnt * kOut * 4);
  ES_HIP(hipMemcpyAsync(n->X, x.data(), x.size() * 4, hipMemcpyHostToDevice, n->stream));
  ES_HIP(hipMemcpyAsync(n->T, t.data(), t.size() * 4, hipMemcpyHostToDevice, n->stream));
  ES_HIP(hipStreamSynchronize(n->stream));
  return ES_OK;
}

}  // namespace

extern "C" {

int es_mlp_create(uint32_t batch_cap, es_mlp** out) {
  if (!out) return fail(ES_ERR_INVALID, "null argument");
  *out = nullptr;
  if (batch_cap == 0 || batch_cap > (1u << 20)) return fail(ES_ERR_INVALID, "batch_cap in [1, 2^20]");
  std::unique_ptr<es_mlp> n(new es_mlp());
  n->cap = pad16(batch_cap);
  ES_HIP(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
  const size_t c = n->cap;
  n->params = n->alloc(kParams);
  n->grads = n->alloc(kParams);
  n->m = n->alloc(kParams);
  n->v = n->alloc(kParams);
  n->X = n->alloc(c * kIn);
  n->T = n->alloc(c * kOut);
  n->H1 = n->alloc(c * kH);
  n->H2 = n->alloc(c * kH);
  n->Y = n->alloc(c * kOutPad);
  n->dY = n->alloc(c * kOutPad);
  n->dH1 = n->alloc(c * kH);
  n->dH2 = n->alloc(c * kH);
  n->part = n->alloc(c);
  for (void* p : n->allocs)
    if (!p) return fail(ES_ERR_DEVICE, "hipMalloc failed");
  if (n->allocs.size() != 13) return fail(ES_ERR_DEVICE, "hipMalloc failed");
  ES_HIP(hipMemsetAsync(n->params, 0, kParams * 4, n->stream));
  ES_HIP(hipMemsetAsync(n->m, 0, kParams * 4, n->stream));
  ES_HIP(hipMemsetAsync(n->v, 0, kParams * 4, n->stream));
  ES_HIP(hipStreamSynchronize(n->stream));
  *out = n.release();
  return ES_OK;
}

int es_mlp_destroy(es_mlp* n) {
  delete n;
  return ES_OK;
}

int es_mlp_set_params(es_mlp* n, const float* params, uint32_t count) {
  if (!n || !params || count != kParams) return fail(ES_ERR_INVALID, "expected ES_MLP_PARAMS floats");
  std::vector<float> p(params, params + kParams);
  for (uint32_t i = 0; i < kH; ++i)  // padded output columns stay zero
    for (uint32_t o = kOut; o < kOutPad; ++o) p[oW3 + i * kOutPad + o] = 0.f;
  for (uint32_t o = kOut; o < kOutPad; ++o) p[ob3 + o] = 0.f;
  ES_HIP(hipMemcpyAsync(n->params, p.data(), kParams * 4, hipMemcpyHostToDevice, n->stream));
  ES_HIP(hipMemsetAsync(n->m, 0, kParams * 4, n->stream));
  ES_HIP(hipMemsetAsync(n->v, 0, kParams * 4, n->stream));
  ES_HIP(hipStreamSynchronize(n->stream));
  n->step = 0;
  return ES_OK;
}

int es_mlp_get_params(es_mlp* n, float* params, uint32_t count) {
  if (!n || !params || count != kParams) return fail(ES_ERR_INVALID, "expected ES_MLP_PARAMS floats");
  ES_HIP(hipMemcpyAsync(params, n->params, kParams * 4, hipMemcpyDeviceToHost, n->stream));
  ES_HIP(hipStreamSynchronize(n->stream));
  return ES_OK;
}

int es_mlp_forward(es_mlp* n, const float* X, uint64_t count, float* Y) {
  if (!n || (count && (!X || !Y))) return fail(ES_ERR_INVALID, "null argument");
  std::vector<float> y;
  for (uint64_t s = 0; s < count; s += n->cap) {
    const uint32_t cnt = (uint32_t)std::min<uint64_t>(n->cap, count - s);
    const uint32_t rows = pad16(cnt);
    int rc = upload_batch(n, X + s * kIn, nullptr, cnt, rows);
    if (rc) return rc;
    if ((rc = forward_rows(n, rows))) return rc;
    y.resize((size_t)rows * kOutPad);
    ES_HIP(hipMemcpyAsync(y.data(), n->Y, y.size() * 4, hipMemcpyDeviceToHost, n->stream));
    ES_HIP(hipStreamSynchronize(n->stream));
    for (uint32_t i = 0; i < cnt; ++i)
      for (uint32_t o = 0; o < kOut; ++o) Y[(s + i) * kOut + o] = y[(size_t)i * kOutPad + o];
  }
  return ES_OK;
}

int es_mlp_gradients(es_mlp* n, const float* X, const float* T, uint32_t count,
                     const float w[2], float* grads, float* loss) {
  if (!n || !X || !T || !w || count == 0) return fail(ES_ERR_INVALID, "null argument");
  if (count > n->cap) return fail(ES_ERR_LIMIT, "batch larger than batch_cap");
  const uint32_t rows = pad16(count);
  int rc = upload_batch(n, X, T, count, rows);
  if (rc) return rc;
  if ((rc = backward_rows(n, count, rows, w[0], w[1], loss))) return rc;
  if (grads) {
    ES_HIP(hipMemcpyAsync(grads, n->grads, kParams * 4, hipMemcpyDeviceToHost, n->stream));
    ES_HIP(hipStreamSynchronize(n->stream));
  }
  return ES_OK;
}

int es_mlp_adam_step(es_mlp* n, const float* X, const float* T, uint32_t count, const float w[2],
                     float lr, float b1, float b2, float eps, float* loss) {
  int rc = es_mlp_gradients(n, X, T, count, w, nullptr, loss);
  if (rc) return rc;
  if ((rc = adam(n, lr, b1, b2, eps))) return rc;
  ES_HIP(hipStreamSynchronize(n->stream));
  return ES_OK;
}

int es_mlp_upload_dataset(es_mlp* n, const float* X, const float* T, uint64_t count) {
  if (!n || !X || !T || count == 0) return fail(ES_ERR_INVALID, "null argument");
  if (count > (1u << 26)) return fail(ES_ERR_LIMIT, "dataset above 2^26 rows");
  hipFree(n->DX);
  hipFree(n->DT);
  hipFree(n->idx);
  n->DX = n->DT = nullptr;
  n->idx = nullptr;
  ES_HIP(hipMalloc((void**)&n->DX, count * kIn * 4));
  ES_HIP(hipMalloc((void**)&n->DT, count * kOut * 4));
  ES_HIP(hipMalloc((void**)&n->idx, count * 4));
  ES_HIP(hipMemcpyAsync(n->DX, X, count * kIn * 4, hipMemcpyHostToDevice, n->stream));
  ES_HIP(hipMemcpyAsync(n->DT, T, count * kOut * 4, hipMemcpyHostToDevice, n->stream));
  ES_HIP(hipStreamSynchronize(n->stream));
  n->n_data = count;
  return ES_OK;
}

int es_mlp_train(es_mlp* n, uint32_t epochs, uint32_t batch, uint64_t seed, const float w[2],
                 float lr, float b1, float b2, float eps, float* last_loss, double* ms) {
  if (!n || !w || batch == 0 || n->n_data == 0) return fail(ES_ERR_INVALID, "no dataset or batch");
  if (batch > n->cap) return fail(ES_ERR_LIMIT, "batch larger than batch_cap");
  const uint64_t N = n->n_data;
  std::vector<uint32_t> perm(N);
  for (uint64_t i = 0; i < N; ++i) perm[i] = (uint32_t)i;
  uint64_t s = seed ? seed : 0x9e3779b97f4a7c15ull;
  hipEvent_t e0, e1;
  ES_HIP(hipEventCreate(&e0));
  ES_HIP(hipEventCreate(&e1));
  ES_HIP(hipEventRecord(e0, n->stream));
  float loss = 0.f;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    for (uint64_t i = N - 1; i > 0; --i) {  // Fisher-Yates with xorshift64*
      s ^= s >> 12;
      s ^= s << 25;
      s ^= s >> 27;
      const uint64_t r = (s * 0x2545F4914F6CDD1Dull) % (i + 1);
      std::swap(perm[i], perm[r]);
    }
    ES_HIP(hipMemcpyAsync(n->idx, perm.data(), N * 4, hipMemcpyHostToDevice, n->stream));
    for (uint64_t b = 0; b < N; b += batch) {
      const uint32_t cnt = (uint32_t)std::min<uint64_t>(batch, N - b);
      const uint32_t rows = pad16(cnt);
      hipLaunchKernelGGL(es_gather_kernel, dim3(rows), dim3(64), 0, n->stream, n->DX, n->DT,
                         n->idx + b, n->X, n->T, cnt, rows);
      ES_HIP(hipGetLastError());
      const bool last = ep + 1 == epochs && b + batch >= N;
      int rc = backward_rows(n, cnt, rows, w[0], w[1], last ? &loss : nullptr);
      if (rc) return rc;
      if ((rc = adam(n, lr, b1, b2, eps))) return rc;
    }
    // the permutation buffer is rewritten next epoch: wait for this epoch's gathers
    ES_HIP(hipStreamSynchronize(n->stream));
  }
  ES_HIP(hipEventRecord(e1, n->stream));
  ES_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  ES_HIP(hipEventElapsedTime(&t, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (last_loss) *last_loss = loss;
  if (ms) *ms = t;
  return ES_OK;
}

}  // extern "C"
