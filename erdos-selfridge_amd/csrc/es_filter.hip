// es_filter.hip -- candidate filter pass over a 128-bit residue stream (HBM-bound).
//
// Input: a stream V of n values (u128, resident in HBM) and a list of u32 candidate IDs into it.
// Output: every ID whose value n passes the Kummer admissibility tables of all filter primes
// (n mod q_j admissible for every j), as the reference's late filters do for the candidates a
// word-mask pass leaves (sieve/es_word_masks.py:87-88 with sieve/gpu_sieve_v8.py:316-322):
// the original ID is what is emitted.
//
// Per candidate: one coalesced 4-byte ID load and one 16-byte value load, then
// n mod q from the value's four 32-bit limbs in FP64 (sum of limb * (2^32k mod q) < 2^52 for
// q < 2^18, so it is exact; quotient by one multiply with 1/q and one fused multiply-subtract, corrected
// by at most one q), and a bit test in the filter's table.  Tables of the most selective
// filters sit in LDS.  Candidates surviving the first filters are packed per wave before the
// rest are tested (es_filter_kernel).  Survivors are compacted with a wave ballot: one atomic
// per wave, each lane writes at its prefix count.
#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "es_common.h"

using es::fail;
using es::mod_u128;

namespace {

// Filter record, 6 doubles: 2^32, 2^64, 2^96 mod q, q, 1/q, then (u32 table word offset, u32
// in-LDS flag).  (Testing a Kummer table by the base-p digits of n mod q instead of reading it
// was measured slower.  So was a residue left in [0, 3q) without the two corrections, read
// from tables holding each bitmap three times: 4.30 ms against 4.16-4.23 at 2^30 candidates,
// the larger tables costing more than the arithmetic saved.)
constexpr int kFiltRecDw = 12;

__device__ __forceinline__ uint32_t fp_mod(double d0, double d1, double d2, double d3,
                                           const double* f) {
  const double acc = fma(d3, f[2], fma(d2, f[1], fma(d1, f[0], d0)));
  const double q = f[3];
  const double qt = floor(acc * f[4]);
  double r = fma(-qt, q, acc);
  r = r < 0.0 ? r + q : r;
  r = r >= q ? r - q : r;
  return (uint32_t)r;
}

// Filter test of one value (limbs as doubles) against filters [j0, j1); with wave_exit the loop
// ends once no lane of the wave is alive.
template <bool WAVE_EXIT>
__device__ __forceinline__ bool es_filters(double d0, double d1, double d2, double d3, bool alive,
                                           uint32_t j0, uint32_t j1, const double* __restrict__ recs,
                                           uint32_t n_lds, const unsigned long long* lbits,
                                           const unsigned long long* __restrict__ gbits) {
  for (uint32_t j = j0; j < j1; ++j) {
    if (WAVE_EXIT && !__any(alive)) break;
    const double* f = recs + kFiltRecDw / 2 * j;
    const uint32_t off = reinterpret_cast<const uint32_t*>(f)[10];
    const uint32_t r = fp_mod(d0, d1, d2, d3, f);
    const unsigned long long wv = j < n_lds ? lbits[off + (r >> 6)] : gbits[off + (r >> 6)];
    alive = alive && ((wv >> (r & 63u)) & 1ull);
  }
  return alive;
}

// Survivors of a wave to the output: one atomic per wave, lanes write at their prefix count.
__device__ __forceinline__ void es_emit(bool alive, uint32_t id, uint32_t* __restrict__ out,
                                        unsigned long long* __restrict__ out_n, uint64_t out_cap) {
  const unsigned long long mask = __ballot(alive);
  if (!mask) return;
  const uint32_t lane = __lane_id();
  const uint32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
  unsigned long long slot0 = 0;
  const uint32_t leader = __ffsll((long long)mask) - 1;
  if (lane == leader) slot0 = atomicAdd(out_n, (unsigned long long)__popcll(mask));
  slot0 = __shfl(slot0, leader);
  if (alive && slot0 + rank < out_cap) out[slot0 + rank] = id;
}

// Two-stage pass.  Stage 1: every lane tests its UNROLL candidates against the first n_first
// filters (no early exit: the lanes are dense).  The few candidates left (~6% after two filters
// at k = 397) are packed into the wave's own 64-entry LDS list (ID and value), and when the
// next batch would overflow it, stage 2 runs the remaining filters over the list with every
// lane holding a live candidate, ending once the wave has none left.  A wave stops after a
// handful of filters instead of testing until the last of 64 sparse lanes dies.  Lists are
// private to a wave (no workgroup barrier); the flush condition is wave-uniform.
template <int UNROLL>
__global__ void __launch_bounds__(256)
es_filter_kernel(const uint32_t* __restrict__ ids, uint64_t n_ids,
                 const uint4* __restrict__ values, const double* __restrict__ recs,
                 uint32_t n_filt, uint32_t n_first, uint32_t n_lds,
                 const unsigned long long* __restrict__ gbits, uint32_t lds_words64,
                 uint32_t* __restrict__ out, unsigned long long* __restrict__ out_n,
                 uint64_t out_cap) {
  extern __shared__ unsigned long long lbits[];
  for (uint32_t z = threadIdx.x; z < lds_words64; z += blockDim.x) lbits[z] = gbits[z];
  __syncthreads();
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  // wave lists after the tables (16-byte aligned): values [nwaves][64], then IDs [nwaves][64]
  uint4* qv_all = reinterpret_cast<uint4*>(lbits + ((lds_words64 + 1u) & ~1u));
  uint4* qv = qv_all + wave * 64u;
  uint32_t* qi = reinterpret_cast<uint32_t*>(qv_all + nwaves * 64u) + wave * 64u;
  uint32_t cnt = 0;  // entries in this wave's list (wave-uniform)
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool have = lane < cnt;
    const uint32_t id = have ? qi[lane] : 0u;
    const uint4 v = have ? qv[lane] : make_uint4(0, 0, 0, 0);
    const bool alive = es_filters<true>((double)v.x, (double)v.y, (double)v.z, (double)v.w, have,
                                        n_first, n_filt, recs, n_lds, lbits, gbits);
    es_emit(alive, id, out, out_n, out_cap);
    cnt = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // round r of a wave covers candidates wb + lane + u * blockDim.x, u < UNROLL, where the
  // wave base wb is wave-uniform, so every lane of a wave runs the same rounds and the list
  // state stays uniform; lanes past n_ids carry no candidate.  Loads are unconditional (a lane
  // past the end reads the last ID and value 0 and is masked), so the code between them is
  // straight-line and each wait covers exactly the loads it needs.
  auto load_ids = [&](uint64_t wb, uint32_t (&id)[UNROLL], bool (&ok)[UNROLL]) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const uint64_t i = wb + lane + (uint64_t)u * blockDim.x;
      ok[u] = i < n_ids;
      id[u] = __builtin_nontemporal_load(ids + (ok[u] ? i : n_ids - 1));
    }
  };
  // values are read once: nontemporal 16-byte loads (7.0-7.1 TB/s in a streaming read on this
  // device against 6.0-6.3 TB/s for cached ones)
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  auto load_vals = [&](const uint32_t (&id)[UNROLL], const bool (&ok)[UNROLL], uint4 (&v)[UNROLL]) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const v4u x = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(values) +
                                               (ok[u] ? id[u] : 0u));
      v[u] = make_uint4(x.x, x.y, x.z, x.w);
    }
  };
  auto stage1 = [&](const uint32_t (&id)[UNROLL], const bool (&ok)[UNROLL],
                    const uint4 (&v)[UNROLL]) {
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const bool alive = es_filters<false>((double)v[u].x, (double)v[u].y, (double)v[u].z,
                                           (double)v[u].w, ok[u], 0, n_first, recs, n_lds, lbits,
                                           gbits);
      const unsigned long long mask = __ballot(alive);
      const uint32_t n = __popcll(mask);
      if (cnt + n > 64u) flush();
      if (alive) {
        const uint32_t slot = cnt + __popcll(mask & ((1ull << lane) - 1ull));
        qi[slot] = id[u];
        qv[slot] = v[u];
      }
      cnt += n;
    }
  };
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * UNROLL;
  const uint64_t wb0 = ((uint64_t)blockIdx.x * blockDim.x) * UNROLL + (threadIdx.x & ~63u);
  // Software pipeline over three slots (IDs, flags, values).  A round r in slot X: filter it,
  // reuse X for the IDs of round r+3, then load the values of round r+2 into slot Z, whose IDs
  // arrived a round ago.  The slots' roles rotate with period 3, written out as three bodies,
  // so no register is copied (a copy of a load's destination would wait for the load).
  // Vector-memory loads retire in order; the IDs of a round are issued before the values of
  // the round before it, so neither wait (values of r, IDs of r+2) drains the values of r+1.
  uint32_t idA[UNROLL], idB[UNROLL], idC[UNROLL];
  bool okA[UNROLL], okB[UNROLL], okC[UNROLL];
  uint4 vA[UNROLL], vB[UNROLL], vC[UNROLL];
  load_ids(wb0, idA, okA);
  load_ids(wb0 + stride, idB, okB);
  load_vals(idA, okA, vA);
  load_ids(wb0 + 2 * stride, idC, okC);
  load_vals(idB, okB, vB);
  auto body = [&](uint64_t wb, uint32_t (&idX)[UNROLL], bool (&okX)[UNROLL],
                  const uint4 (&vX)[UNROLL], const uint32_t (&idZ)[UNROLL],
                  const bool (&okZ)[UNROLL], uint4 (&vZ)[UNROLL]) {
    stage1(idX, okX, vX);
    load_ids(wb + 3 * stride, idX, okX);
    load_vals(idZ, okZ, vZ);
  };
  for (uint64_t wb = wb0; wb < n_ids; wb += 3 * stride) {
    body(wb, idA, okA, vA, idC, okC, vC);
    if (wb + stride >= n_ids) break;
    body(wb + stride, idB, okB, vB, idA, okA, vA);
    if (wb + 2 * stride >= n_ids) break;
    body(wb + 2 * stride, idC, okC, vC, idB, okB, vB);
  }
  if (cnt) flush();
}

// Seeded generator of the value stream: n = (sum_p P_p * ((r_p * inv_p) mod m_p)) mod P + J*P,
// with r_p a random admissible residue mod m_p of each conditioned prime power (P = prod m_p,
// P_p = P / m_p, inv_p = P_p^-1 mod m_p, so each term is the CRT basis times r_p, reduced) and
// J uniform in [0, j_span).  IDs: i with its low 6 bits XORed by a per-64-tile random value (a
// permutation inside each tile of 64, so the value loads of a wave stay in one 1 KiB span).
struct GenPrime {
  uint64_t Pp_lo, Pp_hi;  // P / m_p
  uint32_t m, inv, res_off, res_count;
};

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

__global__ void __launch_bounds__(256)
es_generate_kernel(uint4* __restrict__ values, uint32_t* __restrict__ ids, uint64_t n,
                   const GenPrime* __restrict__ gp, uint32_t n_gp,
                   const uint32_t* __restrict__ res, uint64_t P_lo, uint64_t P_hi,
                   uint64_t j_span, uint64_t seed) {
  const u128 P = es::mk128(P_lo, P_hi);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    u128 s = 0;
    for (uint32_t j = 0; j < n_gp; ++j) {
      const GenPrime g = gp[j];
      const uint64_t rnd = splitmix(seed ^ (i * 0x100000001b3ull) ^ ((uint64_t)j << 56));
      const uint32_t a = res[g.res_off + (uint32_t)(rnd % g.res_count)];
      const uint32_t e = (uint32_t)(((uint64_t)a * g.inv) % g.m);
      s += (u128)e * es::mk128(g.Pp_lo, g.Pp_hi);  // < P
      if (s >= P) s -= P;
    }
    const uint64_t J = splitmix(seed ^ 0x5bd1e995ull ^ (i * 0x9e3779b97f4a7c15ull)) % j_span;
    s += (u128)J * P;
    values[i] = make_uint4((uint32_t)s, (uint32_t)(s >> 32), (uint32_t)(s >> 64),
                           (uint32_t)(s >> 96));
    const uint64_t tile = i >> 6;
    const uint32_t x = (uint32_t)(splitmix(seed * 31ull + tile) & 63ull);
    const uint64_t id = (i & ~63ull) | ((i & 63ull) ^ x);
    ids[i] = id < n ? (uint32_t)id : (uint32_t)i;
  }
}

}  // namespace

struct es_fpass {
  hipStream_t stream = nullptr;
  uint32_t n_filt = 0, n_lds = 0, lds_words64 = 0;
  void* d_recs = nullptr;
  unsigned long long* d_bits = nullptr;
  uint4* d_values = nullptr;
  uint32_t* d_ids = nullptr;
  uint64_t n_values = 0, n_ids = 0;
  uint32_t* d_out = nullptr;
  unsigned long long* d_count = nullptr;
  uint64_t out_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  int cus = 256;
  uint32_t unroll = 2, n_first = 3;  // candidates per lane per round, stage-1 filters
};

extern "C" {

int es_fpass_create(uint32_t n_filters, const uint32_t* filt_q, const uint32_t* filt_off,
                    const uint64_t* words, uint32_t n_words, uint32_t lds_budget,
                    uint64_t out_cap, es_fpass** out) {
  if (!out || (n_filters && (!filt_q || !filt_off || !words)))
    return fail(ES_ERR_INVALID, "null argument");
  *out = nullptr;
  std::unique_ptr<es_fpass> g(new es_fpass());
  es_fpass* f = g.get();
  f->n_filt = n_filters;
  // LDS gets the longest prefix of filters whose tables fit (filters are in test order)
  const uint32_t budget = (lds_budget ? lds_budget : 65536u) / 8;
  std::vector<double> recs((size_t)(kFiltRecDw / 2) * std::max<uint32_t>(n_filters, 1), 0.0);
  std::vector<unsigned long long> lds_img, glob_img;
  uint32_t nl = 0;
  for (uint32_t j = 0; j < n_filters; ++j) {
    const uint32_t q = filt_q[j];
    // q < 2^18 keeps the FP64 limb sum below 2^52 (exact)
    if (q < 2 || q >= (1u << 18)) return fail(ES_ERR_LIMIT, "filter modulus outside [2, 2^18)");
    const uint32_t nw = (q + 63) / 64;
    if ((uint64_t)filt_off[j] + nw > n_words) return fail(ES_ERR_INVALID, "bitmap outside words");
    const bool to_lds = (j == nl) && lds_img.size() + nw <= budget;
    std::vector<unsigned long long>& img = to_lds ? lds_img : glob_img;
    const uint32_t off = (uint32_t)img.size();
    img.insert(img.end(), words + filt_off[j], words + filt_off[j] + nw);
    if (to_lds) ++nl;
    double* r = &recs[(size_t)(kFiltRecDw / 2) * j];
    r[0] = (double)mod_u128((u128)1 << 32, q);
    r[1] = (double)mod_u128((u128)1 << 64, q);
    r[2] = (double)mod_u128((u128)1 << 96, q);
    r[3] = (double)q;
    r[4] = 1.0 / (double)q;
    reinterpret_cast<uint32_t*>(r)[10] = off;  // rebased below for global
    reinterpret_cast<uint32_t*>(r)[11] = to_lds ? 1u : 0u;
  }
  // one device array: LDS image first (copied into LDS per workgroup), then global tables
  const uint32_t lds_n = (uint32_t)lds_img.size();
  for (uint32_t j = 0; j < n_filters; ++j) {
    uint32_t* rr = reinterpret_cast<uint32_t*>(&recs[(size_t)(kFiltRecDw / 2) * j]);
    if (!rr[11]) rr[10] += lds_n;
  }
  std::vector<unsigned long long> all(lds_img);
  all.insert(all.end(), glob_img.begin(), glob_img.end());
  if (all.empty()) all.push_back(0);
  f->n_lds = nl;
  f->lds_words64 = lds_n;
  f->out_cap = out_cap ? out_cap : (1ull << 24);
  ES_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
  ES_HIP(hipMalloc(&f->d_recs, recs.size() * 8));
  ES_HIP(hipMemcpy(f->d_recs, recs.data(), recs.size() * 8, hipMemcpyHostToDevice));
  ES_HIP(hipMalloc((void**)&f->d_bits, all.size() * 8));
  ES_HIP(hipMemcpy(f->d_bits, all.data(), all.size() * 8, hipMemcpyHostToDevice));
  ES_HIP(hipMalloc((void**)&f->d_out, f->out_cap * 4));
  ES_HIP(hipMalloc((void**)&f->d_count, 8));
  ES_HIP(hipEventCreate(&f->ev0));
  ES_HIP(hipEventCreate(&f->ev1));
  int dev = 0;
  ES_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  ES_HIP(hipGetDeviceProperties(&prop, dev));
  f->cus = prop.multiProcessorCount;
  *out = g.release();
  return ES_OK;
}

int es_fpass_destroy(es_fpass* f) {
  if (!f) return ES_OK;
  if (f->stream) hipStreamSynchronize(f->stream);
  hipFree(f->d_recs);
  hipFree(f->d_bits);
  hipFree(f->d_values);
  hipFree(f->d_ids);
  hipFree(f->d_out);
  hipFree(f->d_count);
  if (f->ev0) hipEventDestroy(f->ev0);
  if (f->ev1) hipEventDestroy(f->ev1);
  if (f->stream) hipStreamDestroy(f->stream);
  delete f;
  return ES_OK;
}

static int fpass_alloc(es_fpass* f, uint64_t n_values, uint64_t n_ids) {
  if (f->n_values != n_values) {
    if (f->d_values) ES_HIP(hipFree(f->d_values));
    f->d_values = nullptr;
    ES_HIP(hipMalloc((void**)&f->d_values, std::max<uint64_t>(n_values, 1) * 16));
    f->n_values = n_values;
  }
  if (f->n_ids != n_ids) {
    if (f->d_ids) ES_HIP(hipFree(f->d_ids));
    f->d_ids = nullptr;
    ES_HIP(hipMalloc((void**)&f->d_ids, std::max<uint64_t>(n_ids, 1) * 4));
    f->n_ids = n_ids;
  }
  return ES_OK;
}

int es_fpass_upload(es_fpass* f, const uint64_t* values_lohi, uint64_t n_values,
                    const uint32_t* ids, uint64_t n_ids) {
  if (!f || (n_values && !values_lohi) || (n_ids && !ids)) return fail(ES_ERR_INVALID, "null argument");
  for (uint64_t i = 0; i < n_ids; ++i)
    if (ids[i] >= n_values) return fail(ES_ERR_INVALID, "candidate ID outside the stream");
  int rc = fpass_alloc(f, n_values, n_ids);
  if (rc) return rc;
  if (n_values) ES_HIP(hipMemcpy(f->d_values, values_lohi, n_values * 16, hipMemcpyHostToDevice));
  if (n_ids) ES_HIP(hipMemcpy(f->d_ids, ids, n_ids * 4, hipMemcpyHostToDevice));
  return ES_OK;
}

int es_fpass_generate(es_fpass* f, uint64_t n, uint64_t seed, uint32_t n_mods,
                      const uint32_t* mods, const uint32_t* res_count, const uint32_t* residues,
                      uint64_t j_span) {
  if (!f || (n_mods && (!mods || !res_count || !residues))) return fail(ES_ERR_INVALID, "null argument");
  if (n == 0 || n > (1ull << 32)) return fail(ES_ERR_INVALID, "stream length must be in [1, 2^32]");
  if (j_span == 0) return fail(ES_ERR_INVALID, "multiplier span must be positive");
  u128 P = 1;
  for (uint32_t j = 0; j < n_mods; ++j) {
    if (mods[j] < 2) return fail(ES_ERR_INVALID, "modulus must be >= 2");
    P *= mods[j];
    if ((P >> 112) != 0) return fail(ES_ERR_INVALID, "generator modulus must stay below 2^112");
  }
  if (((P * (u128)j_span) >> 127) != 0) return fail(ES_ERR_INVALID, "values would exceed 2^127");
  std::vector<GenPrime> gp(std::max<uint32_t>(n_mods, 1));
  uint32_t total = 0;
  for (uint32_t j = 0; j < n_mods; ++j) {
    const uint32_t m = mods[j];
    const u128 Pp = P / m;
    // inverse of Pp mod m by extended Euclid (moduli pairwise coprime)
    int64_t a = (int64_t)(Pp % m), b = m, x0 = 1, x1 = 0;
    while (b) {
      const int64_t q = a / b, t = a - q * b;
      a = b;
      b = t;
      const int64_t tx = x0 - q * x1;
      x0 = x1;
      x1 = tx;
    }
    if (a != 1) return fail(ES_ERR_INVALID, "generator moduli must be pairwise coprime");
    if (res_count[j] == 0) return fail(ES_ERR_INVALID, "empty residue list");
    for (uint32_t r = 0; r < res_count[j]; ++r)
      if (residues[total + r] >= m) return fail(ES_ERR_INVALID, "residue not reduced");
    gp[j].Pp_lo = (uint64_t)Pp;
    gp[j].Pp_hi = (uint64_t)(Pp >> 64);
    gp[j].m = m;
    gp[j].inv = (uint32_t)(((x0 % (int64_t)m) + m) % m);
    gp[j].res_off = total;
    gp[j].res_count = res_count[j];
    total += res_count[j];
  }
  int rc = fpass_alloc(f, n, n);
  if (rc) return rc;
  void *d_gp = nullptr, *d_res = nullptr;
  ES_HIP(hipMalloc(&d_gp, gp.size() * sizeof(GenPrime)));
  ES_HIP(hipMalloc(&d_res, std::max<uint32_t>(total, 1) * 4));
  ES_HIP(hipMemcpy(d_gp, gp.data(), gp.size() * sizeof(GenPrime), hipMemcpyHostToDevice));
  if (total) ES_HIP(hipMemcpy(d_res, residues, total * 4, hipMemcpyHostToDevice));
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, (uint64_t)f->cus * 64);
  hipLaunchKernelGGL(es_generate_kernel, dim3((uint32_t)blocks), dim3(256), 0, f->stream,
                     f->d_values, f->d_ids, n, (const GenPrime*)d_gp, n_mods,
                     (const uint32_t*)d_res, (uint64_t)P, (uint64_t)(P >> 64), j_span, seed);
  ES_HIP(hipGetLastError());
  ES_HIP(hipStreamSynchronize(f->stream));
  ES_HIP(hipFree(d_gp));
  ES_HIP(hipFree(d_res));
  return ES_OK;
}

int es_fpass_set_shape(es_fpass* f, uint32_t unroll, uint32_t n_first) {
  if (!f) return fail(ES_ERR_INVALID, "null argument");
  if (unroll != 2 && unroll != 4 && unroll != 8) return fail(ES_ERR_INVALID, "unroll must be 2, 4 or 8");
  f->unroll = unroll;
  f->n_first = n_first;
  return ES_OK;
}

int es_fpass_run(es_fpass* f, uint64_t* n_out) {
  if (!f) return fail(ES_ERR_INVALID, "null argument");
  if (f->n_ids == 0) {  // nothing to test (the kernel's loads assume at least one ID)
    f->last_ms = 0;
    if (n_out) *n_out = 0;
    return ES_OK;
  }
  ES_HIP(hipMemsetAsync(f->d_count, 0, 8, f->stream));
  const uint32_t U = f->unroll, threads = 256, nwaves = threads / 64;
  const uint64_t per_wg = (uint64_t)threads * U;
  uint64_t blocks = (f->n_ids + per_wg - 1) / per_wg;
  blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)f->cus * 32));
  const uint32_t n_first = std::min(f->n_first, f->n_filt);
  const size_t lds = (size_t)((f->lds_words64 + 1u) & ~1u) * 8 + (size_t)nwaves * 64 * 20;
  auto kern = U == 8 ? es_filter_kernel<8> : U == 4 ? es_filter_kernel<4> : es_filter_kernel<2>;
  if (lds > 65536)
    ES_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  ES_HIP(hipEventRecord(f->ev0, f->stream));
  hipLaunchKernelGGL(kern, dim3((uint32_t)blocks), dim3(threads), lds, f->stream, f->d_ids,
                     f->n_ids, f->d_values, (const double*)f->d_recs, f->n_filt, n_first,
                     f->n_lds, f->d_bits, f->lds_words64, f->d_out, f->d_count, f->out_cap);
  ES_HIP(hipGetLastError());
  ES_HIP(hipEventRecord(f->ev1, f->stream));
  unsigned long long c = 0;
  ES_HIP(hipMemcpyAsync(&c, f->d_count, 8, hipMemcpyDeviceToHost, f->stream));
  ES_HIP(hipStreamSynchronize(f->stream));
  float ms = 0.f;
  ES_HIP(hipEventElapsedTime(&ms, f->ev0, f->ev1));
  f->last_ms = ms;
  if (n_out) *n_out = c;
  if (c > f->out_cap) return fail(ES_ERR_OVERFLOW, "survivor buffer overflow");
  return ES_OK;
}

int es_fpass_read(es_fpass* f, uint32_t* ids, uint64_t count) {
  if (!f || (count && !ids)) return fail(ES_ERR_INVALID, "null argument");
  if (count > f->out_cap) return fail(ES_ERR_INVALID, "read past the survivor buffer");
  if (count) ES_HIP(hipMemcpy(ids, f->d_out, count * 4, hipMemcpyDeviceToHost));
  return ES_OK;
}

int es_fpass_read_stream(es_fpass* f, uint64_t offset, uint64_t count, uint64_t* values_lohi,
                         uint32_t* ids) {
  if (!f) return fail(ES_ERR_INVALID, "null argument");
  if (offset > f->n_values || count > f->n_values - offset)
    return fail(ES_ERR_INVALID, "read outside the stream");
  if (values_lohi && count)
    ES_HIP(hipMemcpy(values_lohi, f->d_values + offset, count * 16, hipMemcpyDeviceToHost));
  if (ids && count) {
    if (offset > f->n_ids || count > f->n_ids - offset) return fail(ES_ERR_INVALID, "read outside IDs");
    ES_HIP(hipMemcpy(ids, f->d_ids + offset, count * 4, hipMemcpyDeviceToHost));
  }
  return ES_OK;
}

int es_fpass_last_ms(es_fpass* f, double* ms) {
  if (!f || !ms) return fail(ES_ERR_INVALID, "null argument");
  *ms = f->last_ms;
  return ES_OK;
}

}  // extern "C"
