// es_mitm.hip -- bounded n-window enumeration by meet in the middle (sieve/gpu_mitm.py:32-125).
//
// The wheel's two halves: half 1 (decoded per index in the kernel) and half 2 (a table of all
// its CRT partial sums mod M, built and sorted on the device: one decode kernel, then two
// stable 64-bit radix sorts, low word then high word, giving 128-bit order).  For a window
// [a, b) of s = n - t*M, the half-2 sums meeting s1 lie on the circular run
// [(a - s1) mod M, +(b - a)) of the sorted table: two lower-bound searches give its ends.
// Every n formed is tested against the filter tables (n mod q from 128-bit limbs and
// (2^32k mod q), plus t*M mod q), with early exit; survivors are appended.
//
// Lanes: one lane per s1 when runs are short (most lanes only search), or one wave per s1 when
// runs are long (the wave walks the run 64 entries at a time, coalesced table reads).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "es_common.h"
#include "../../include/es_mitm.h"

using es::fail;
using es::mod_u128;

namespace {

constexpr int kFiltRec = 8;  // q, bitmap word offset, 2^32 / 2^64 / 2^96 mod q, t*M mod q, pad

struct MitmArgs {
  uint64_t M_lo, M_hi, a_lo, a_hi, len_lo, len_hi;  // window [a, a + len) of s, len <= M
  uint32_t w1, nf;
  const uint32_t* radix1;
  const uint32_t* off1;
  const uint64_t* t1lo;
  const uint64_t* t1hi;
  uint64_t n2;
  const uint64_t* s2hi;
  const uint64_t* s2lo;
  const uint32_t* frec;
  const double* fdrec;           // per filter: 2^32, 2^64, 2^96 mod q, q, 1/q, t*M mod q
  uint32_t n_first;              // filters every candidate is tested against before packing
  const uint32_t* fbits;
  unsigned long long* out;       // survivors: lo, hi pairs
  unsigned long long* out_n;
  unsigned long long* tested;    // candidates formed
  uint64_t out_cap;
};

__device__ __forceinline__ u128 mitm_s1(const MitmArgs& A, uint64_t idx, u128 M) {
  u128 s = 0;
  uint64_t rem = idx;
  for (uint32_t j = 0; j < A.w1; ++j) {
    const uint32_t r = A.radix1[j];
    uint32_t d;
    if (j + 1 == A.w1) {
      d = (uint32_t)rem;
    } else {
      const uint64_t q = rem / r;
      d = (uint32_t)(rem - q * r);
      rem = q;
    }
    const uint32_t o = A.off1[j] + d;
    s += es::mk128(A.t1lo[o], A.t1hi[o]);
    if (s >= M) s -= M;
  }
  return s;
}

// first index u in [0, n2) with table[u] >= x (128-bit order)
__device__ __forceinline__ uint64_t mitm_lower_bound(const MitmArgs& A, u128 x) {
  uint64_t lo = 0, hi = A.n2;
  const uint64_t xh = (uint64_t)(x >> 64), xl = (uint64_t)x;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    const uint64_t h = A.s2hi[mid], l = A.s2lo[mid];
    if (h < xh || (h == xh && l < xl)) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool mitm_test(const MitmArgs& A, u128 s) {
  const uint64_t lo = (uint64_t)s, hi = (uint64_t)(s >> 64);
  for (uint32_t j = 0; j < A.nf; ++j) {
    cu32_t* f = (cu32_t*)A.frec + kFiltRec * j;  // uniform record: scalar loads
    const uint32_t q = f[0];
    uint32_t r = es_mod128(lo, hi, q, f[2], f[3], f[4]) + f[5];
    if (r >= q) r -= q;
    if (!((A.fbits[f[1] + (r >> 5)] >> (r & 31u)) & 1u)) return false;
  }
  return true;
}

// Filters [j0, j1) in FP64 (every q < 2^18: limb * (2^32k mod q) < 2^50, the sum of the four
// terms and t*M mod q stays below 2^52 and is exact; the quotient from one multiply by 1/q is
// corrected by one q either way).  EXIT 0: every filter (dense lanes); 1: the loop ends once no
// lane of the wave is alive; 2: each lane stops at its first failing filter.
template <int EXIT>
__device__ __forceinline__ bool mitm_test_fp(const MitmArgs& A, u128 s, bool alive, uint32_t j0,
                                             uint32_t j1) {
  const double d0 = (double)(uint32_t)s, d1 = (double)(uint32_t)(s >> 32),
               d2 = (double)(uint32_t)(s >> 64), d3 = (double)(uint32_t)(s >> 96);
  for (uint32_t j = j0; j < j1; ++j) {
    if (EXIT == 1 && !__any(alive)) break;
    if (EXIT == 2 && !alive) break;
    cf64_t* f = (cf64_t*)A.fdrec + 8 * j;  // uniform record: scalar loads
    const double acc = fma(d3, f[2], fma(d2, f[1], fma(d1, f[0], d0 + f[5])));
    const double q = f[3];
    double r = fma(-floor(acc * f[4]), q, acc);
    r = r < 0.0 ? r + q : r;
    r = r >= q ? r - q : r;
    const uint32_t ri = (uint32_t)r;
    const uint32_t off = ((cu32_t*)A.frec)[kFiltRec * j + 1];
    alive = alive && ((A.fbits[off + (ri >> 5)] >> (ri & 31u)) & 1u);
  }
  return alive;
}

__device__ __forceinline__ void mitm_emit(const MitmArgs& A, u128 s) {
  const unsigned long long slot = atomicAdd(A.out_n, 1ull);
  if (slot < A.out_cap) {
    A.out[2 * slot] = (uint64_t)s;
    A.out[2 * slot + 1] = (uint64_t)(s >> 64);
  }
}

// The circular run of table indices meeting s1: [r0s, r0e) and [r1s, r1e).
__device__ __forceinline__ void mitm_run(const MitmArgs& A, u128 s1, u128 M, uint64_t& r0s,
                                         uint64_t& r0e, uint64_t& r1s, uint64_t& r1e) {
  const u128 a = es::mk128(A.a_lo, A.a_hi), len = es::mk128(A.len_lo, A.len_hi);
  const u128 L = a >= s1 ? a - s1 : a + (M - s1);  // (a - s1) mod M
  const u128 E = L + len;                          // < 2M
  r0s = mitm_lower_bound(A, L);
  r1s = r1e = 0;
  if (E <= M) {
    r0e = E == M ? A.n2 : mitm_lower_bound(A, E);
  } else {
    r0e = A.n2;
    r1e = mitm_lower_bound(A, E - M);
  }
}

template <bool WAVE>
__global__ void __launch_bounds__(256) es_mitm_kernel(MitmArgs A, uint64_t i0, uint64_t count) {
  const u128 M = es::mk128(A.M_lo, A.M_hi);
  unsigned long long tested = 0;
  if (!WAVE) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += stride) {
      const u128 s1 = mitm_s1(A, i0 + t, M);
      uint64_t r0s, r0e, r1s, r1e;
      mitm_run(A, s1, M, r0s, r0e, r1s, r1e);
      for (int part = 0; part < 2; ++part) {
        const uint64_t st = part ? r1s : r0s, en = part ? r1e : r0e;
        for (uint64_t u = st; u < en; ++u) {
          u128 s = s1 + es::mk128(A.s2lo[u], A.s2hi[u]);
          if (s >= M) s -= M;
          ++tested;
          if (A.fdrec ? mitm_test_fp<2>(A, s, true, 0, A.nf) : mitm_test(A, s))
            mitm_emit(A, s);
        }
      }
    }
  } else {
    // One wave per s1: every lane decodes and searches the same s1 (uniform, broadcast loads),
    // then the wave walks the run with lane stride 64.  With FP64 filters: two stages -- every
    // candidate against the first n_first filters, the few left packed into the wave's LDS
    // list, the remaining filters run over the full list (lanes dense, early exit per wave).
    __shared__ uint4 q_all[4 * 64];  // 256 threads = 4 waves x 64 entries (s as 4 limbs)
    const uint32_t lane = __lane_id();
    uint4* qv = q_all + (threadIdx.x >> 6) * 64u;
    const bool fp = A.fdrec != nullptr;
    uint32_t cnt = 0;  // wave-uniform
    auto flush = [&]() {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const bool have = lane < cnt;
      const uint4 v = have ? qv[lane] : make_uint4(0, 0, 0, 0);
      const u128 s = es::mk128(((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z);
      if (mitm_test_fp<1>(A, s, have, A.n_first, A.nf)) mitm_emit(A, s);
      cnt = 0;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t t = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < count;
         t += waves) {
      const u128 s1 = mitm_s1(A, i0 + t, M);
      uint64_t r0s, r0e, r1s, r1e;
      mitm_run(A, s1, M, r0s, r0e, r1s, r1e);
      for (int part = 0; part < 2; ++part) {
        const uint64_t st = part ? r1s : r0s, en = part ? r1e : r0e;
        for (uint64_t u0 = st; u0 < en; u0 += 64) {  // wave-uniform bound
          const uint64_t u = u0 + lane;
          const bool ok = u < en;
          u128 s = 0;
          if (ok) {
            s = s1 + es::mk128(A.s2lo[u], A.s2hi[u]);
            if (s >= M) s -= M;
            ++tested;
          }
          if (!fp) {
            if (ok && mitm_test(A, s)) mitm_emit(A, s);
            continue;
          }
          const bool alive = mitm_test_fp<0>(A, s, ok, 0, A.n_first);
          const unsigned long long mask = __ballot(alive);
          const uint32_t n = __popcll(mask);
          if (cnt + n > 64u) flush();
          if (alive) {
            const uint32_t slot = cnt + __popcll(mask & ((1ull << lane) - 1ull));
            qv[slot] = make_uint4((uint32_t)s, (uint32_t)(s >> 32), (uint32_t)(s >> 64),
                                  (uint32_t)(s >> 96));
          }
          cnt += n;
        }
      }
    }
    if (fp && cnt) flush();
  }
  if (tested) atomicAdd(A.tested, tested);
}

// Sorted sweep: half 1 is a sorted table too.  A wave takes `chunk` consecutive s1 of it and
// its 64/GL groups of GL lanes step through them in lockstep: at each step group g takes the
// g-th of the next 64/GL consecutive s1, so the runs of one step nearly coincide and the wave
// reads one small, cache-resident region of the half-2 table.  Each group walks its s1's run
// with lane stride GL (the walk loop runs to the longest run of the step, wave-uniform).  As
// s1 grows, (a - s1) mod M and (a + len - s1) mod M fall, so a group's two run ends move down
// the table by short steps, which its GL lanes take GL entries at a time; a binary search is
// needed only at the chunk's first step and where an end wraps past 0 or M.
constexpr int kWalkUnroll = 2;

template <int GL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
es_mitm_sweep_kernel(MitmArgs A, const uint64_t* __restrict__ s1hi,
                     const uint64_t* __restrict__ s1lo, uint64_t n1, uint64_t chunk,
                     uint64_t c_lo, uint64_t c_hi) {
  constexpr uint32_t NG = 64 / GL;
  const u128 M = es::mk128(A.M_lo, A.M_hi);
  const u128 a = es::mk128(A.a_lo, A.a_hi), len = es::mk128(A.len_lo, A.len_hi);
  __shared__ uint4 q_all[4 * 64];
  const uint32_t lane = __lane_id(), gl = lane % GL, grp = lane / GL;
  uint4* qv = q_all + (threadIdx.x >> 6) * 64u;
  const bool fp = A.fdrec != nullptr;
  unsigned long long tested = 0;
  uint32_t cnt = 0;  // wave-uniform
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool have = lane < cnt;
    const uint4 v = have ? qv[lane] : make_uint4(0, 0, 0, 0);
    const u128 s = es::mk128(((uint64_t)v.y << 32) | v.x, ((uint64_t)v.w << 32) | v.z);
    if (mitm_test_fp<1>(A, s, have, A.n_first, A.nf)) mitm_emit(A, s);
    cnt = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  auto ge = [&](uint64_t u, u128 x) {  // table[u] >= x
    const uint64_t h = A.s2hi[u], l = A.s2lo[u];
    const uint64_t xh = (uint64_t)(x >> 64), xl = (uint64_t)x;
    return h > xh || (h == xh && l >= xl);
  };
  // Lowers p (a lower bound of x in the table that x has since fallen below) to the lower
  // bound of x: the group's lanes test GL entries below p per round (wave-uniform rounds).
  auto lower = [&](uint64_t& p, u128 x, bool moving) {
    const uint64_t gmask = GL == 64 ? ~0ull : ((1ull << GL) - 1ull);
    while (__ballot(moving) != 0) {
      const bool c = moving && p > gl && ge(p - 1 - gl, x);
      const uint64_t bits = (__ballot(c) >> (grp * GL)) & gmask;
      if (moving) {
        // entries >= x just below p (a prefix of the GL tested, the table being sorted)
        const uint32_t m = bits == gmask ? (uint32_t)GL : (uint32_t)__builtin_ctzll(~bits);
        p -= m;
        moving = m == GL;
      }
    }
  };
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  for (uint64_t c = c_lo + (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       c < c_hi && c * chunk < n1; c += waves) {  // wave-uniform
    const uint64_t w_lo = c * chunk, w_hi = n1 - w_lo < chunk ? n1 : w_lo + chunk;
    const uint64_t steps = (w_hi - w_lo + NG - 1) / NG;
    uint64_t pX = 0, pY = 0;
    u128 Xp = 0, Tp = 0;
    bool wp = false;
    for (uint64_t step = 0; step < steps; ++step) {  // wave-uniform
      const uint64_t i = w_lo + step * NG + grp;
      const bool have = i < w_hi;
      u128 s1 = 0, X = 0, T = 0;
      bool w = false, searchX = false, searchY = false;
      if (have) {
        s1 = es::mk128(s1lo[i], s1hi[i]);
        X = a >= s1 ? a - s1 : a + (M - s1);  // (a - s1) mod M
        const u128 Y = X + len;               // < 2M
        w = Y > M;
        T = w ? Y - M : Y;                    // end of the run's last part, <= M
        searchX = step == 0 || X > Xp;
        searchY = step == 0 || w != wp || T > Tp;
        if (searchX) pX = mitm_lower_bound(A, X);
        if (searchY) pY = T == M ? A.n2 : mitm_lower_bound(A, T);
      }
      lower(pX, X, have && !searchX);
      lower(pY, T, have && !searchY);
      uint64_t st = 0, L = 0;  // run: L table entries from st, circular
      if (have) {
        Xp = X;
        Tp = T;
        wp = w;
        st = pX;
        L = w ? (A.n2 - pX) + pY : pY - pX;
      }
      // kWalkUnroll rounds of the walk per trip: every round's table loads are issued before
      // the first round is tested, so their latencies overlap
      constexpr int U = kWalkUnroll;
      for (uint64_t k0 = 0; __ballot(k0 < L) != 0; k0 += (uint64_t)GL * U) {  // wave-uniform
        uint64_t tl[U], th[U];
        bool okv[U];
#pragma unroll
        for (int r = 0; r < U; ++r) {
          const uint64_t k = k0 + (uint64_t)r * GL + gl;
          okv[r] = k < L;
          tl[r] = th[r] = 0;
          if (okv[r]) {
            uint64_t u = st + k;
            if (u >= A.n2) u -= A.n2;
            tl[r] = A.s2lo[u];
            th[r] = A.s2hi[u];
          }
        }
#pragma unroll
        for (int r = 0; r < U; ++r) {
          const bool ok = okv[r];
          u128 s = 0;
          if (ok) {
            s = s1 + es::mk128(tl[r], th[r]);
            if (s >= M) s -= M;
            ++tested;
          }
          if (!fp) {
            if (ok && mitm_test(A, s)) mitm_emit(A, s);
            continue;
          }
          const bool alive = mitm_test_fp<0>(A, s, ok, 0, A.n_first);
          const unsigned long long mask = __ballot(alive);
          const uint32_t n = __popcll(mask);
          if (cnt + n > 64u) flush();
          if (alive) {
            const uint32_t slot = cnt + __popcll(mask & ((1ull << lane) - 1ull));
            qv[slot] = make_uint4((uint32_t)s, (uint32_t)(s >> 32), (uint32_t)(s >> 64),
                                  (uint32_t)(s >> 96));
          }
          cnt += n;
        }
      }
    }
  }
  if (fp && cnt) flush();
  if (tested) atomicAdd(A.tested, tested);
}

// table[i] = CRT partial sum of half-2 index i (mixed radix, ring 0 fastest)
__global__ void __launch_bounds__(256)
es_mitm_table_kernel(uint32_t w2, const uint32_t* __restrict__ radix2,
                     const uint32_t* __restrict__ off2, const uint64_t* __restrict__ t2lo,
                     const uint64_t* __restrict__ t2hi, uint64_t M_lo, uint64_t M_hi, uint64_t n2,
                     uint64_t* __restrict__ hi_out, uint64_t* __restrict__ lo_out) {
  const u128 M = es::mk128(M_lo, M_hi);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2;
       i += (uint64_t)gridDim.x * blockDim.x) {
    u128 s = 0;
    uint64_t rem = i;
    for (uint32_t j = 0; j < w2; ++j) {
      const uint32_t r = radix2[j];
      const uint32_t d = (uint32_t)(rem % r);
      rem /= r;
      const uint32_t o = off2[j] + d;
      s += es::mk128(t2lo[o], t2hi[o]);
      if (s >= M) s -= M;
    }
    hi_out[i] = (uint64_t)(s >> 64);
    lo_out[i] = (uint64_t)s;
  }
}

}  // namespace

struct es_mitm {
  hipStream_t stream = nullptr;
  u128 M = 0;
  uint64_t n1 = 1, n2 = 1;
  uint32_t w1 = 0, nf = 0, wave = 0;
  std::vector<uint32_t> filt_q, frec;
  std::vector<double> fdrec;  // FP64 filter records when every q < 2^18
  double* d_fdrec = nullptr;
  uint32_t n_first = 2;
  std::vector<void*> allocs;
  uint32_t *d_radix1 = nullptr, *d_off1 = nullptr, *d_frec = nullptr, *d_fbits = nullptr;
  uint64_t *d_t1lo = nullptr, *d_t1hi = nullptr, *d_s2hi = nullptr, *d_s2lo = nullptr;
  uint64_t *d_s1hi = nullptr, *d_s1lo = nullptr;  // sorted half 1 (sweep), or null
  unsigned long long *d_out = nullptr, *d_cnt = nullptr;  // d_cnt[0] survivors, [1] tested
  uint64_t out_cap = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  uint64_t last_tested = 0;
  int cus = 256;
  ~es_mitm() {
    if (stream) hipStreamSynchronize(stream);
    for (void* p : allocs) hipFree(p);
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    if (stream) hipStreamDestroy(stream);
  }
  template <typename T>
  int alloc(T** p, size_t n) {
    void* q = nullptr;
    ES_HIP(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    allocs.push_back(q);
    *p = (T*)q;
    return ES_OK;
  }
};

extern "C" {

int es_mitm_create(const es_mitm_desc* d, es_mitm** out) {
  if (!d || !out) return fail(ES_ERR_INVALID, "null argument");
  *out = nullptr;
  if (d->abi_version != ES_ABI_VERSION) return fail(ES_ERR_INVALID, "ABI version mismatch");
  const u128 M = es::mk128(d->M_lo, d->M_hi);
  if (M == 0 || (M >> 126) != 0) return fail(ES_ERR_INVALID, "wheel modulus must be in [1, 2^126)");
  if (d->n1 < 1 || d->n2 < 1 || !d->radix1 || !d->radix2 || !d->term1_lo || !d->term1_hi ||
      !d->term2_lo || !d->term2_hi)
    return fail(ES_ERR_INVALID, "each half needs at least one ring (a unit ring for an empty half)");
  std::unique_ptr<es_mitm> g(new es_mitm());
  es_mitm* m = g.get();
  m->M = M;
  m->w1 = d->n1;
  if (d->walk > 2 && d->walk != 4 && d->walk != 8 && d->walk != 16 && d->walk != 32 &&
      d->walk != 64)
    return fail(ES_ERR_INVALID, "walk must be 0, 1, 2, 4, 8, 16, 32 or 64");
  m->wave = d->walk;
  // half 1: radices, offsets, terms
  std::vector<uint32_t> off1(d->n1);
  uint64_t t1 = 0;
  for (uint32_t j = 0; j < d->n1; ++j) {
    if (d->radix1[j] == 0) return fail(ES_ERR_INVALID, "empty ring");
    off1[j] = (uint32_t)t1;
    t1 += d->radix1[j];
    m->n1 *= d->radix1[j];
    if (m->n1 >= (1ull << 62)) return fail(ES_ERR_LIMIT, "half 1 above 2^62 entries");
  }
  std::vector<uint32_t> off2(d->n2);
  uint64_t t2 = 0;
  for (uint32_t j = 0; j < d->n2; ++j) {
    if (d->radix2[j] == 0) return fail(ES_ERR_INVALID, "empty ring");
    off2[j] = (uint32_t)t2;
    t2 += d->radix2[j];
    m->n2 *= d->radix2[j];
    if (m->n2 >= (1ull << 32)) return fail(ES_ERR_LIMIT, "half 2 (the table) above 2^32 entries");
  }
  for (uint64_t i = 0; i < t1; ++i)
    if (es::mk128(d->term1_lo[i], d->term1_hi[i]) >= M) return fail(ES_ERR_INVALID, "term not reduced mod M");
  for (uint64_t i = 0; i < t2; ++i)
    if (es::mk128(d->term2_lo[i], d->term2_hi[i]) >= M) return fail(ES_ERR_INVALID, "term not reduced mod M");
  // filters
  m->nf = d->n_filters;
  if (m->nf && (!d->filt_q || !d->filt_off || !d->filt_bits)) return fail(ES_ERR_INVALID, "filter arrays missing");
  m->frec.assign((size_t)kFiltRec * std::max<uint32_t>(m->nf, 1), 0);
  for (uint32_t f = 0; f < m->nf; ++f) {
    const uint32_t q = d->filt_q[f];
    if (q < 2 || q >= (1u << 24)) return fail(ES_ERR_LIMIT, "filter modulus outside [2, 2^24)");
    if ((uint64_t)d->filt_off[f] + (q + 31) / 32 > d->n_filt_words)
      return fail(ES_ERR_INVALID, "filter bitmap outside filt_bits");
    uint32_t* r = &m->frec[(size_t)kFiltRec * f];
    r[0] = q;
    r[1] = d->filt_off[f];
    r[2] = mod_u128((u128)1 << 32, q);
    r[3] = mod_u128((u128)1 << 64, q);
    r[4] = mod_u128((u128)1 << 96, q);
    r[5] = 0;  // t*M mod q, per scan
    m->filt_q.push_back(q);
  }
  bool fp_ok = m->nf > 0;
  for (uint32_t f = 0; f < m->nf; ++f) fp_ok = fp_ok && d->filt_q[f] < (1u << 18);
  if (fp_ok) {
    m->fdrec.assign((size_t)8 * m->nf, 0.0);
    for (uint32_t f = 0; f < m->nf; ++f) {
      const uint32_t* r = &m->frec[(size_t)kFiltRec * f];
      double* x = &m->fdrec[(size_t)8 * f];
      x[0] = r[2];
      x[1] = r[3];
      x[2] = r[4];
      x[3] = (double)r[0];
      x[4] = 1.0 / (double)r[0];
      x[5] = 0.0;  // t*M mod q, per scan
    }
  }
  m->n_first = std::min<uint32_t>(2, m->nf);
  m->out_cap = d->out_cap ? d->out_cap : (1ull << 20);
  ES_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  ES_HIP(hipEventCreate(&m->ev0));
  ES_HIP(hipEventCreate(&m->ev1));
  int dev = 0;
  ES_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  ES_HIP(hipGetDeviceProperties(&prop, dev));
  m->cus = prop.multiProcessorCount;
  int rc;
  uint32_t *d_radix2 = nullptr, *d_off2 = nullptr;
  uint64_t *d_t2lo = nullptr, *d_t2hi = nullptr;
  if ((rc = m->alloc(&m->d_radix1, d->n1)) || (rc = m->alloc(&m->d_off1, d->n1)) ||
      (rc = m->alloc(&m->d_t1lo, t1)) || (rc = m->alloc(&m->d_t1hi, t1)) ||
      (rc = m->alloc(&d_radix2, d->n2)) || (rc = m->alloc(&d_off2, d->n2)) ||
      (rc = m->alloc(&d_t2lo, t2)) || (rc = m->alloc(&d_t2hi, t2)) ||
      (rc = m->alloc(&m->d_frec, m->frec.size())) ||
      (rc = m->alloc(&m->d_fdrec, std::max<size_t>(m->fdrec.size(), 1))) ||
      (rc = m->alloc(&m->d_fbits, std::max<uint32_t>(d->n_filt_words, 1))) ||
      (rc = m->alloc(&m->d_s2hi, m->n2)) || (rc = m->alloc(&m->d_s2lo, m->n2)) ||
      (rc = m->alloc(&m->d_out, 2 * m->out_cap)) || (rc = m->alloc(&m->d_cnt, 2)))
    return rc;
  ES_HIP(hipMemcpy(m->d_radix1, d->radix1, d->n1 * 4, hipMemcpyHostToDevice));
  ES_HIP(hipMemcpy(m->d_off1, off1.data(), d->n1 * 4, hipMemcpyHostToDevice));
  ES_HIP(hipMemcpy(m->d_t1lo, d->term1_lo, t1 * 8, hipMemcpyHostToDevice));
  ES_HIP(hipMemcpy(m->d_t1hi, d->term1_hi, t1 * 8, hipMemcpyHostToDevice));
  ES_HIP(hipMemcpy(d_radix2, d->radix2, d->n2 * 4, hipMemcpyHostToDevice));
  ES_HIP(hipMemcpy(d_off2, off2.data(), d->n2 * 4, hipMemcpyHostToDevice));
  ES_HIP(hipMemcpy(d_t2lo, d->term2_lo, t2 * 8, hipMemcpyHostToDevice));
  ES_HIP(hipMemcpy(d_t2hi, d->term2_hi, t2 * 8, hipMemcpyHostToDevice));
  if (m->nf)
    ES_HIP(hipMemcpy(m->d_fbits, d->filt_bits, (size_t)d->n_filt_words * 4, hipMemcpyHostToDevice));
  // the half-2 table (and, for the sweep, the half-1 table): decode into the table, then sort
  // by low word into scratch and (stably) by high word back into the table; each pass sorts
  // only the bits a value below M can have; sort scratch is freed once the table is built
  const uint32_t hi_bits = d->M_hi ? 64u - (uint32_t)__builtin_clzll(d->M_hi) : 0u;
  const uint32_t lo_bits = d->M_hi ? 64u : 64u - (uint32_t)__builtin_clzll(d->M_lo | 1u);
  auto build_table = [&](uint32_t w, const uint32_t* d_radix, const uint32_t* d_off,
                         const uint64_t* d_tlo, const uint64_t* d_thi, uint64_t n,
                         uint64_t* out_hi, uint64_t* out_lo) -> int {
    uint64_t *k_hi = nullptr, *k_lo = nullptr;
    void* tmp = nullptr;
    int r = ES_OK;
    do {
      if (hipMalloc((void**)&k_hi, n * 8) != hipSuccess || hipMalloc((void**)&k_lo, n * 8) != hipSuccess) {
        r = fail(ES_ERR_DEVICE, "hipMalloc failed for the sort buffers");
        break;
      }
      const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, (uint64_t)m->cus * 64);
      hipLaunchKernelGGL(es_mitm_table_kernel, dim3((uint32_t)blocks), dim3(256), 0, m->stream,
                         w, d_radix, d_off, d_tlo, d_thi, d->M_lo, d->M_hi, n, out_hi, out_lo);
      if (hipGetLastError() != hipSuccess) { r = fail(ES_ERR_DEVICE, "table kernel launch failed"); break; }
      size_t tmp_bytes = 0;
      if (rocprim::radix_sort_pairs(nullptr, tmp_bytes, out_lo, k_lo, out_hi, k_hi, (uint32_t)n,
                                    0, 64, m->stream) != hipSuccess ||
          hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 16)) != hipSuccess) {
        r = fail(ES_ERR_DEVICE, "radix sort set-up failed");
        break;
      }
      // low word first, then the high word (LSD radix sort is stable): 128-bit order
      bool ok = rocprim::radix_sort_pairs(tmp, tmp_bytes, out_lo, k_lo, out_hi, k_hi,
                                          (uint32_t)n, 0, lo_bits, m->stream) == hipSuccess;
      if (ok && hi_bits) {
        ok = rocprim::radix_sort_pairs(tmp, tmp_bytes, k_hi, out_hi, k_lo, out_lo, (uint32_t)n, 0,
                                       hi_bits, m->stream) == hipSuccess;
      } else if (ok) {  // every high word is 0: the low-word order is the table
        ok = hipMemcpyAsync(out_hi, k_hi, n * 8, hipMemcpyDeviceToDevice, m->stream) == hipSuccess &&
             hipMemcpyAsync(out_lo, k_lo, n * 8, hipMemcpyDeviceToDevice, m->stream) == hipSuccess;
      }
      if (!ok || hipStreamSynchronize(m->stream) != hipSuccess) {
        r = fail(ES_ERR_DEVICE, "radix sort failed");
        break;
      }
    } while (false);
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(k_hi);
    (void)hipFree(k_lo);
    (void)hipFree(tmp);
    return r;
  };
  if ((rc = build_table(d->n2, d_radix2, d_off2, d_t2lo, d_t2hi, m->n2, m->d_s2hi, m->d_s2lo)))
    return rc;
  // sweep: half 1 as a sorted table too, when it and its sort buffers fit beside the rest
  bool sweep = d->sweep == 2;
  if (d->sweep == 0 && m->n1 < (1ull << 32)) {
    size_t free_b = 0, total_b = 0;
    ES_HIP(hipMemGetInfo(&free_b, &total_b));
    const long double need = (long double)m->n1 * 40.0L + (long double)(1ull << 30);
    sweep = need < (long double)free_b * 0.8L;
  }
  if (sweep) {
    if (m->n1 >= (1ull << 32)) return fail(ES_ERR_LIMIT, "sorted half 1 above 2^32 entries");
    if ((rc = m->alloc(&m->d_s1hi, m->n1)) || (rc = m->alloc(&m->d_s1lo, m->n1))) return rc;
    if ((rc = build_table(d->n1, m->d_radix1, m->d_off1, m->d_t1lo, m->d_t1hi, m->n1, m->d_s1hi,
                          m->d_s1lo)))
      return rc;
  }
  *out = g.release();
  return ES_OK;
}

int es_mitm_destroy(es_mitm* m) {
  delete m;
  return ES_OK;
}

int es_mitm_sizes(es_mitm* m, uint64_t* n1, uint64_t* n2, uint32_t* swept) {
  if (!m) return fail(ES_ERR_INVALID, "null argument");
  if (n1) *n1 = m->n1;
  if (n2) *n2 = m->n2;
  if (swept) *swept = m->d_s1hi ? 1u : 0u;
  return ES_OK;
}

int es_mitm_scan(es_mitm* m, uint64_t t, uint64_t a_lo, uint64_t a_hi, uint64_t b_lo,
                 uint64_t b_hi, uint64_t* count) {
  if (!m) return fail(ES_ERR_INVALID, "null argument");
  const u128 a = es::mk128(a_lo, a_hi), b = es::mk128(b_lo, b_hi);
  if (!(a < b) || b > m->M) return fail(ES_ERR_INVALID, "window must satisfy 0 <= a < b <= M");
  const u128 len = b - a;
  // block constants t*M mod q
  for (uint32_t f = 0; f < m->nf; ++f) {
    const uint32_t q = m->filt_q[f];
    m->frec[(size_t)kFiltRec * f + 5] = (uint32_t)(((uint64_t)(t % q) * mod_u128(m->M, q)) % q);
    if (!m->fdrec.empty()) m->fdrec[(size_t)8 * f + 5] = (double)m->frec[(size_t)kFiltRec * f + 5];
  }
  ES_HIP(hipMemcpyAsync(m->d_frec, m->frec.data(), m->frec.size() * 4, hipMemcpyHostToDevice,
                        m->stream));
  if (!m->fdrec.empty())
    ES_HIP(hipMemcpyAsync(m->d_fdrec, m->fdrec.data(), m->fdrec.size() * 8,
                          hipMemcpyHostToDevice, m->stream));
  ES_HIP(hipMemsetAsync(m->d_cnt, 0, 16, m->stream));
  MitmArgs A{};
  A.M_lo = (uint64_t)m->M;
  A.M_hi = (uint64_t)(m->M >> 64);
  A.a_lo = a_lo;
  A.a_hi = a_hi;
  A.len_lo = (uint64_t)len;
  A.len_hi = (uint64_t)(len >> 64);
  A.w1 = m->w1;
  A.nf = m->nf;
  A.radix1 = m->d_radix1;
  A.off1 = m->d_off1;
  A.t1lo = m->d_t1lo;
  A.t1hi = m->d_t1hi;
  A.n2 = m->n2;
  A.s2hi = m->d_s2hi;
  A.s2lo = m->d_s2lo;
  A.frec = m->d_frec;
  A.fdrec = m->fdrec.empty() ? nullptr : m->d_fdrec;
  A.n_first = m->n_first;
  A.fbits = m->d_fbits;
  A.out = m->d_out;
  A.out_n = m->d_cnt;
  A.tested = m->d_cnt + 1;
  A.out_cap = m->out_cap;
  // expected run length n2 * len / M decides the walk when not fixed, and bounds each launch
  // to ~2^34 candidates (long scans become many launches of bounded length)
  const long double run = (long double)m->n2 * (long double)len / (long double)m->M;
  bool wave = m->wave == 2 || ((m->wave == 0 || m->wave >= 4) && run >= 8.0L);
  const long double per = (long double)(1ull << 34) / (run > 1.0L ? run : 1.0L);
  const uint64_t chunk = (uint64_t)std::min<long double>(std::max<long double>(per, 65536.0L),
                                                         (long double)(1ull << 26));
  ES_HIP(hipEventRecord(m->ev0, m->stream));
  if (m->d_s1hi) {
    // sorted sweep: 1024 consecutive s1 per wave, 64 / GL per step (GL lanes walk one s1's run:
    // 4 while runs stay below 256 entries -- more concurrent s1 per wave beat shorter walks --
    // or `walk` when it names a size); launches of ~2^34 expected candidates
    uint32_t GL = m->wave >= 4 ? m->wave : 4;
    if (m->wave < 4)
      while (GL < 64 && (long double)GL * 64.0L <= run) GL *= 2;
    const uint64_t per_wave = 1024, chunks = (m->n1 + per_wave - 1) / per_wave;
    const long double cand_chunk = run * (long double)per_wave;
    const uint64_t per_launch = (uint64_t)std::max<long double>(
        1.0L, std::min<long double>((long double)(1ull << 34) / std::max<long double>(cand_chunk, 1.0L),
                                    (long double)chunks));
    for (uint64_t c = 0; c < chunks; c += per_launch) {
      const uint64_t c_hi = std::min<uint64_t>(chunks, c + per_launch);
      const uint64_t blocks = std::max<uint64_t>(
          1, std::min<uint64_t>((c_hi - c + 3) / 4, (uint64_t)m->cus * 64));
      auto kern = GL == 64 ? es_mitm_sweep_kernel<64>
                  : GL == 32 ? es_mitm_sweep_kernel<32>
                  : GL == 16 ? es_mitm_sweep_kernel<16>
                  : GL == 8 ? es_mitm_sweep_kernel<8> : es_mitm_sweep_kernel<4>;
      hipLaunchKernelGGL(kern, dim3((uint32_t)blocks), dim3(256), 0, m->stream, A,
                         (const uint64_t*)m->d_s1hi, (const uint64_t*)m->d_s1lo, m->n1, per_wave,
                         c, c_hi);
      ES_HIP(hipGetLastError());
    }
  }
  for (uint64_t s = 0; s < m->n1 && !m->d_s1hi; s += chunk) {
    const uint64_t cnt = std::min<uint64_t>(chunk, m->n1 - s);
    const uint64_t per_block = wave ? 4 : 256;
    const uint64_t blocks =
        std::max<uint64_t>(1, std::min<uint64_t>((cnt + per_block - 1) / per_block, (uint64_t)m->cus * 64));
    if (wave)
      hipLaunchKernelGGL(es_mitm_kernel<true>, dim3((uint32_t)blocks), dim3(256), 0, m->stream,
                         A, s, cnt);
    else
      hipLaunchKernelGGL(es_mitm_kernel<false>, dim3((uint32_t)blocks), dim3(256), 0, m->stream,
                         A, s, cnt);
    ES_HIP(hipGetLastError());
  }
  ES_HIP(hipEventRecord(m->ev1, m->stream));
  unsigned long long c[2] = {0, 0};
  ES_HIP(hipMemcpyAsync(c, m->d_cnt, 16, hipMemcpyDeviceToHost, m->stream));
  ES_HIP(hipStreamSynchronize(m->stream));
  float ms = 0.f;
  ES_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  m->last_ms = ms;
  m->last_tested = c[1];
  if (count) *count = c[0];
  if (c[0] > m->out_cap) return fail(ES_ERR_OVERFLOW, "survivor buffer overflow");
  return ES_OK;
}

int es_mitm_read(es_mitm* m, uint64_t* s_lohi, uint64_t count) {
  if (!m || (count && !s_lohi)) return fail(ES_ERR_INVALID, "null argument");
  if (count > m->out_cap) return fail(ES_ERR_INVALID, "read past the survivor buffer");
  if (count) ES_HIP(hipMemcpy(s_lohi, m->d_out, count * 16, hipMemcpyDeviceToHost));
  return ES_OK;
}

int es_mitm_last(es_mitm* m, double* ms, uint64_t* candidates) {
  if (!m) return fail(ES_ERR_INVALID, "null argument");
  if (ms) *ms = m->last_ms;
  if (candidates) *candidates = m->last_tested;
  return ES_OK;
}

}  // extern "C"
