// es_sieve_dev.h -- device side of the CRT-wheel sieve engine (es_sieve.hip): plan records,
// the sieve and batch kernels, and their instance selectors.  The kernel instances are compiled
// in four translation units (es_sieve_k0..3.hip, by mask count F) so they build in parallel;
// es_sieve.hip holds the host side and picks an instance through es_dev::pick_kernel.
#pragma once

#include "es_common.h"

namespace es_dev {

constexpr int kMaxMasks = 12;
constexpr int kMaskInfo = 8;   // dwords per mask record
constexpr int kSlowInfo = 16;  // dwords per exact-filter record

// Read only by the exact re-test of queued candidates, through one pointer.
struct ColdPlan {
  uint64_t M0_lo, M0_hi;
  uint32_t q0, r0, n_slow, stride;
  uint32_t slow_fp, pad;      // 1: every exact-filter modulus < 2^18 (FP64 residues are exact)
  const uint32_t* strip;      // strip rows; L_sigma at dwords 2..5
  const uint32_t* e0;         // [r0] ring-0 e values, ascending
  const uint32_t* inner_ids;  // [r0] idx offset of each sorted slot: its ring-0 digit x lead
  const uint32_t* slow;       // [n_slow][16] q, word offset, 2^32, 2^64, 2^96 mod q, t*M mod q,
                              // 2 unused, then 2^32, 2^64, 2^96 mod q and 1/q as doubles
  const uint32_t* slow_bits;  // admissibility bitmaps
  unsigned long long* out;    // survivor IDs
  unsigned long long* out_n;  // [0] survivors appended (may pass out_cap), [1] 1 + the slot of
                              // the least residue s >= [2] recorded so far (0: none), [2] the
                              // least s that counts for the minimum (k + 2 in block 0, else 0)
  unsigned long long* out_s;  // [out_cap][2] residue s (lo, hi) of each survivor slot
  uint64_t out_cap;
};

__device__ __forceinline__ uint64_t es_readlane64(uint64_t v, uint32_t lane) {
  return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)lane) |
         ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)lane) << 32);
}

// Survivor append and running minimum, called by every lane of a pass of the exact test (ok:
// this lane's position passed).  The survivors of the pass take consecutive slots from one
// atomicAdd per wave (ballot prefix), and each slot records its residue s beside its ID.  The
// least s among them that counts for the minimum (s >= out_n[2]) is found across the wave and
// offered once to the running minimum out_n[1], which names the slot of the least s so far: a
// compare-and-swap replaces it only while the offered s is below the named record.  Records
// are written before a slot is published (fence) and read after it is seen (acquire), so the
// word always names a least recorded s.  Survivors beyond out_cap are counted but not
// recorded (the caller sees the overflow).
__device__ __forceinline__ void es_emit(const ColdPlan* __restrict__ C, bool ok, uint64_t idx,
                                        uint64_t s_lo, uint64_t s_hi) {
  const uint64_t okm = __ballot(ok);
  if (okm == 0) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t leader = (uint32_t)__builtin_ctzll(okm);
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(C->out_n, (unsigned long long)__popcll(okm));
  base = es_readlane64(base, leader);
  const uint64_t slot = base + __builtin_amdgcn_mbcnt_hi(
                                   (uint32_t)(okm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)okm, 0u));
  bool cand = false;
  if (ok && slot < C->out_cap) {
    C->out[slot] = idx;
    C->out_s[2 * slot] = s_lo;
    C->out_s[2 * slot + 1] = s_hi;
    cand = s_hi != 0 || s_lo >= C->out_n[2];
  }
  uint64_t cm = __ballot(cand);
  if (cm == 0) return;
  const uint32_t first = (uint32_t)__builtin_ctzll(cm);
  uint64_t bl = ~0ull, bh = ~0ull, bslot = 0;
  while (cm) {  // wave-uniform: the least s among the candidate lanes
    const uint32_t src = (uint32_t)__builtin_ctzll(cm);
    cm &= cm - 1;
    const uint64_t l = es_readlane64(s_lo, src), h = es_readlane64(s_hi, src);
    if (h < bh || (h == bh && l < bl)) {
      bl = l;
      bh = h;
      bslot = es_readlane64(slot, src);
    }
  }
  __threadfence();  // the wave's records are visible before any slot is published
  if (lane != first) return;
  unsigned long long* best = C->out_n + 1;
  unsigned long long cur = __hip_atomic_load(best, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    if (cur) {
      unsigned long long* r = C->out_s + 2 * (cur - 1);
      const uint64_t rl = __hip_atomic_load(r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t rh = __hip_atomic_load(r + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (rh < bh || (rh == bh && rl <= bl)) break;  // the recorded minimum stands
    }
    unsigned long long want = cur;
    if (__hip_atomic_compare_exchange_strong(best, &want, (unsigned long long)(bslot + 1),
                                             __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                             __HIP_MEMORY_SCOPE_AGENT))
      break;
    cur = want;
  }
}

// Kernel argument.
struct HotPlan {
  uint64_t M0_lo, M0_hi;
  uint32_t q0, r0, S, nb, W, stride;
  uint32_t lds_words, carry_base;  // LDS image size (words), carry table byte offset
  const uint32_t* lds_src;         // device copy of the LDS image
  const uint32_t* strip;           // [S][stride] H, X, L(4), 0, 0, Q8_f (row byte step per f);
                                   // X = idx correction of a lead block
  const uint32_t* base_radix;      // [nb]
  const uint32_t* base_off;        // [nb] first term of each base ring
  const uint32_t* term_h;          // base-ring terms, h part
  const uint64_t* term_l;          // base-ring terms, l part (lo, hi pairs)
  const uint32_t* masks;           // [F][8] m, table byte base, M0 mod m, 2^32, 2^64, 2^96 mod m,
                                   //        row multiplier (-M)^-1 mod m, t*M mod m
  const uint32_t* gmask;           // mask tables served from global memory (through L1)
  const ColdPlan* cold;
};

// Exact re-test of the candidate slots in `bits` (sorted ring-0 slots 32*w + b) of outer
// index o (strip row sg); h, kap are the outer index's ring-0 row and l-carry, z = M0 - l_u.
// CF & 2: a lead block (the kernel's ring 0 is the reference's ring j, its rings 1..j the
// reference's rings 0..j-1), and `o` is passed as the outer index's idx base, o * r0 + the
// strip row's signed correction (dword 1 of the row): a slot's idx is that + its offset.
// Kernels with one or two words per row inline it into the queue drain (an out-of-line call
// makes the whole kernel allocate registers around the call: 96 VGPRs and spills against 79
// and none inline for 6 masks, one word); wider rows, whose unrolled word loop has one drain
// per word, call it out of line to keep the code small.
// The core takes the outer index's residue b = h * M0 + l (h < q0, l < M0) directly.
template <int CF>
__device__ __forceinline__ void es_exact_test_hl(const ColdPlan* __restrict__ C, uint32_t bits,
                                              uint32_t w, uint32_t h, u128 l, uint64_t o,
                                              uint64_t idx_lo, uint64_t idx_hi) {
  const u128 M0 = es::mk128(C->M0_lo, C->M0_hi);
  const uint32_t q0 = C->q0, r0 = C->r0, ns = C->n_slow;
  const bool fp = C->slow_fp != 0;
  while (bits) {
    const uint32_t i = 32u * w + (uint32_t)__builtin_ctz(bits);
    bits &= bits - 1u;
    const uint64_t idx = (CF & 2) ? o + C->inner_ids[i] : o * r0 + C->inner_ids[i];
    // a position outside the window takes part in the pass's append as a non-survivor
    bool ok = idx >= idx_lo && idx < idx_hi;
    uint32_t hh = h + C->e0[i];
    if (hh >= q0) hh -= q0;
    const u128 s = (u128)hh * M0 + l;
    const uint64_t s_lo = (uint64_t)s, s_hi = (uint64_t)(s >> 64);
    // s stays live as its FP64 limbs only (each below 2^32, so exact): the record of a
    // survivor is rebuilt from them after the test
    const double d0 = (double)(uint32_t)s_lo, d1 = (double)(uint32_t)(s_lo >> 32),
                 d2 = (double)(uint32_t)s_hi, d3 = (double)(uint32_t)(s_hi >> 32);
    for (uint32_t j = 0; j < ns && ok; ++j) {
      cu32_t* f = (cu32_t*)C->slow + kSlowInfo * j;  // same j in every active lane: scalar
      const uint32_t q = f[0];
      uint32_t r;
      if (fp) {
        // every term below 2^50 and the sum below 2^53: exact; the quotient from one multiply
        // by 1/q is off by at most one either way
        cf64_t* fd = reinterpret_cast<cf64_t*>(f + 8);
        const double qd = (double)q;
        const double acc = fma(d3, fd[2], fma(d2, fd[1], fma(d1, fd[0], d0 + (double)f[5])));
        double x = fma(-floor(acc * fd[3]), qd, acc);
        x = x < 0.0 ? x + qd : x;
        x = x >= qd ? x - qd : x;
        r = (uint32_t)x;
      } else {
        const uint64_t lo = (uint64_t)(uint32_t)d0 | ((uint64_t)(uint32_t)d1 << 32);
        const uint64_t hi = (uint64_t)(uint32_t)d2 | ((uint64_t)(uint32_t)d3 << 32);
        r = es_mod128(lo, hi, q, f[2], f[3], f[4]) + f[5];
        if (r >= q) r -= q;
      }
      ok = (C->slow_bits[f[1] + (r >> 5)] >> (r & 31u)) & 1u;
    }
    es_emit(C, ok, idx, (uint64_t)(uint32_t)d0 | ((uint64_t)(uint32_t)d1 << 32),
            (uint64_t)(uint32_t)d2 | ((uint64_t)(uint32_t)d3 << 32));
  }
}

template <int CF>
__device__ __forceinline__ void es_exact_test(const ColdPlan* __restrict__ C, uint32_t bits,
                                           uint32_t w, uint32_t h, uint32_t kap, uint64_t zl,
                                           uint64_t zh, uint32_t sg, uint64_t o,
                                           uint64_t idx_lo, uint64_t idx_hi) {
  const uint32_t* row = C->strip + (size_t)sg * C->stride;
  const u128 L = es::mk128((uint64_t)row[2] | ((uint64_t)row[3] << 32),
                       (uint64_t)row[4] | ((uint64_t)row[5] << 32));
  const u128 M0 = es::mk128(C->M0_lo, C->M0_hi);
  const u128 z = es::mk128(zl, zh);
  u128 l = (M0 - z) + L;  // l_u + L_sigma
  if (kap) l -= M0;
  es_exact_test_hl<CF>(C, bits, w, h, l, o, idx_lo, idx_hi);
}

static __device__ __noinline__ void es_exact_test_call(const ColdPlan* __restrict__ C, uint32_t bits,
                                                uint32_t w, uint32_t h, uint32_t kap,
                                                uint64_t zl, uint64_t zh, uint32_t sg,
                                                uint64_t o, uint64_t idx_lo, uint64_t idx_hi) {
  es_exact_test<0>(C, bits, w, h, kap, zl, zh, sg, o, idx_lo, idx_hi);
}

// Per-wave candidate queue in LDS, after the tables (kWaveQueueBytes per wave): up to
// kQueueCap entries {bits, (sg * words + w) | lane << 26} of mask words that still have
// candidates.  Entries wait until 64 are queued and are then re-tested exactly, one per lane,
// so the exact test runs on full waves; a strip's leftovers are re-tested when it ends.  The
// entry's lane supplies its strip data (z, outer base, h_u) through cross-lane permutes.
constexpr uint32_t kQueueCap = 128;
constexpr uint32_t kWaveQueueBytes = kQueueCap * 8;
constexpr uint32_t kQueueIndexBits = 26;  // sg * words + w < 2^26 (checked at engine creation)

__device__ __forceinline__ uint2* es_wave_queue(const HotPlan& P, const uint32_t* lds) {
  const uint32_t base = (P.lds_words * 4u + 15u) & ~15u;
  return reinterpret_cast<uint2*>(reinterpret_cast<char*>(const_cast<uint32_t*>(lds)) + base +
                                  (threadIdx.x >> 6) * kWaveQueueBytes);
}

__device__ __forceinline__ uint32_t es_from_lane(uint32_t src, uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}

// Exact re-test of queued entries [first, first + n), one per lane.  Called by the whole wave
// (the permutes read every lane's strip data).
template <int W, int C>
__device__ __forceinline__ void es_drain(const HotPlan& P, const uint2* q, uint32_t first,
                                         uint32_t n, uint32_t hu, uint64_t zl, uint64_t zh,
                                         uint64_t obase, uint64_t idx_lo, uint64_t idx_hi) {
  __builtin_amdgcn_wave_barrier();
  const uint32_t lane = threadIdx.x & 63u;
  const uint2 ent = lane < n ? q[first + lane] : make_uint2(0u, 0u);
  const uint32_t src = ent.y >> kQueueIndexBits;
  const uint64_t szl = (uint64_t)es_from_lane(src, (uint32_t)zl) |
                       ((uint64_t)es_from_lane(src, (uint32_t)(zl >> 32)) << 32);
  const uint64_t szh = (uint64_t)es_from_lane(src, (uint32_t)zh) |
                       ((uint64_t)es_from_lane(src, (uint32_t)(zh >> 32)) << 32);
  const uint64_t sob = (uint64_t)es_from_lane(src, (uint32_t)obase) |
                       ((uint64_t)es_from_lane(src, (uint32_t)(obase >> 32)) << 32);
  const uint32_t shu = es_from_lane(src, hu);
  if (lane < n) {
    const uint32_t sw = ent.y & ((1u << kQueueIndexBits) - 1u);
    const uint32_t nw = W ? (uint32_t)W : P.W;
    const uint32_t sg = W == 1 ? sw : sw / nw, w = sw - sg * nw;
    const uint32_t* row = P.strip + (size_t)sg * P.stride;
    const uint64_t Ll = (uint64_t)row[2] | ((uint64_t)row[3] << 32);
    const uint64_t Lh = (uint64_t)row[4] | ((uint64_t)row[5] << 32);
    const bool kap = es_ge128(Lh, Ll, szh, szl);
    uint32_t h = shu + row[0] + (kap ? 1u : 0u);
    if (h >= P.q0) h -= P.q0;
    // lead block: the idx base of the outer index (see es_exact_test)
    const uint64_t ob = (C & 2) ? (sob + sg) * P.r0 + (uint64_t)(int64_t)(int32_t)row[1]
                                : sob + sg;
    if (W == 1 || W == 2)
      es_exact_test<C>(P.cold, ent.x, w, h, kap, szl, szh, sg, ob, idx_lo, idx_hi);
    else
      es_exact_test_call(P.cold, ent.x, w, h, kap, szl, szh, sg, sob + sg, idx_lo, idx_hi);
  }
  __builtin_amdgcn_wave_barrier();
}

// Queues this lane's candidate word (bits != 0); drains 64 entries once that many wait.
// `cnt` is wave-uniform.
template <int W, int C>
__device__ __forceinline__ void es_enqueue(const HotPlan& P, uint2* q, uint32_t& cnt,
                                           uint32_t bits, uint32_t sw, uint32_t hu, uint64_t zl,
                                           uint64_t zh, uint64_t obase, uint64_t idx_lo,
                                           uint64_t idx_hi) {
  const uint64_t any = __ballot(bits != 0u);
  if (any == 0) return;
  if (bits) {
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(
        (uint32_t)(any >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)any, 0u));
    q[cnt + below] = make_uint2(bits, sw | ((threadIdx.x & 63u) << kQueueIndexBits));
  }
  cnt += (uint32_t)__popcll(any);
  if (cnt >= 64u) {
    cnt -= 64u;
    es_drain<W, C>(P, q, cnt, 64u, hu, zl, zh, obase, idx_lo, idx_hi);
  }
}

// One strip: S outer indices, uniform across the wave.  W = words per ring-0 mask (0 = runtime).
// The last G of the F masks are read from global memory (vector loads through L1) instead of
// LDS, so the two load pipes share the row traffic.  A[f] is the byte address of the lane's
// row in table f; an outer sum that wraps past M starts one row on.  The words for slots
// without an inner wrap and with one are ANDed over the masks first and merged by the carry
// word once.
template <int F, int G, int W, int C, bool CHECK>
__device__ __forceinline__ void es_sweep(const HotPlan& P, const uint32_t* lds,
                                         const uint32_t (&A)[F + 1], uint32_t hu, uint64_t zl,
                                         uint64_t zh, uint64_t obase, uint32_t sb, uint32_t se,
                                         uint2* wq, uint64_t idx_lo, uint64_t idx_hi) {
  const uint32_t q0 = P.q0, S = P.S, stride = P.stride;
  uint32_t cnt = 0;
  const uint32_t nw = W ? (uint32_t)W : P.W;
  const char* lb = reinterpret_cast<const char*>(lds);
  const char* gb = reinterpret_cast<const char*>(P.gmask);
  const uint32_t* carry_tab = reinterpret_cast<const uint32_t*>(lb + P.carry_base);
  for (uint32_t sg = 0; sg < S; ++sg) {
    cu32_t* row = (cu32_t*)(P.strip) + (size_t)sg * stride;
    const uint32_t Hs = row[0];
    const uint64_t Ll = (uint64_t)row[2] | ((uint64_t)row[3] << 32);
    const uint64_t Lh = (uint64_t)row[4] | ((uint64_t)row[5] << 32);
    const bool kap = es_ge128(Lh, Ll, zh, zl);      // l_u + L_sigma >= M0
    uint32_t h = hu + Hs + (kap ? 1u : 0u);
    const bool c1 = h >= q0;                        // b_u + B_sigma >= M
    h = c1 ? h - q0 : h;
    const bool live = !CHECK || (sg >= sb && sg < se);
    const uint32_t wrap = c1 ? 8u * nw : 0u;        // one row on
    if (W == 1) {
      uint32_t carry;
      if (C & 1) {  // byte table of j(h): bits i >= j carry
        const uint32_t j = reinterpret_cast<const uint8_t*>(carry_tab)[h];
        carry = j >= 32u ? 0u : 0xffffffffu << j;
      } else {
        carry = carry_tab[h];
      }
      uint32_t bx = 0xffffffffu, by = 0xffffffffu;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const uint32_t a = A[f] + row[8 + f] + wrap;
        const uint2 v = f < F - G ? *reinterpret_cast<const uint2*>(lb + a)
                                  : *reinterpret_cast<const uint2*>(gb + a);
        bx &= v.x;
        by &= v.y;
      }
      uint32_t bits = (carry & by) | (~carry & bx);
      if (F == 0) bits = P.r0 >= 32 ? 0xffffffffu : ((1u << P.r0) - 1u);
      if (!live) bits = 0;
      es_enqueue<W, C>(P, wq, cnt, bits, sg, hu, zl, zh, obase, idx_lo, idx_hi);
    } else if (W == 2) {
      const uint2 carry = *reinterpret_cast<const uint2*>(carry_tab + 2 * h);
      uint32_t b0x = 0xffffffffu, b0y = 0xffffffffu, b1x = 0xffffffffu, b1y = 0xffffffffu;
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const uint32_t a = A[f] + row[8 + f] + wrap;
        const uint4 v = *reinterpret_cast<const uint4*>(lb + a);
        b0x &= v.x;
        b0y &= v.y;
        b1x &= v.z;
        b1y &= v.w;
      }
      uint32_t b0 = (carry.x & b0y) | (~carry.x & b0x);
      uint32_t b1 = (carry.y & b1y) | (~carry.y & b1x);
      if (F == 0) b1 = P.r0 >= 64 ? 0xffffffffu : ((1u << (P.r0 - 32)) - 1u);
      if (!live) b0 = b1 = 0;
      es_enqueue<W, C>(P, wq, cnt, b0, 2 * sg, hu, zl, zh, obase, idx_lo, idx_hi);
      es_enqueue<W, C>(P, wq, cnt, b1, 2 * sg + 1, hu, zl, zh, obase, idx_lo, idx_hi);
    } else {
      // carrying slots are the suffix [j, r0) of the sorted ring-0 slots
      const uint32_t j = reinterpret_cast<const uint16_t*>(carry_tab)[h];
#pragma unroll
      for (uint32_t w = 0; w < nw; ++w) {
        const int d = (int)j - 32 * (int)w;
        const uint32_t carry = d <= 0 ? 0xffffffffu : (d >= 32 ? 0u : 0xffffffffu << d);
        uint32_t bx = 0xffffffffu, by = 0xffffffffu;
#pragma unroll
        for (int f = 0; f < F; ++f) {
          const uint32_t a = A[f] + row[8 + f] + wrap + 8u * w;
          const uint2 v = *reinterpret_cast<const uint2*>(lb + a);
          bx &= v.x;
          by &= v.y;
        }
        uint32_t bits = (carry & by) | (~carry & bx);
        if (F == 0) {
          const uint32_t n = P.r0 - 32u * w;
          bits = n >= 32u ? 0xffffffffu : ((1u << n) - 1u);
        }
        if (!live) bits = 0;
        es_enqueue<W, C>(P, wq, cnt, bits, sg * nw + w, hu, zl, zh, obase, idx_lo, idx_hi);
      }
    }
  }
  if (cnt) es_drain<W, C>(P, wq, 0, cnt, hu, zl, zh, obase, idx_lo, idx_hi);
}

// Strip u of a plan (u may be a replay of a valid strip for lanes past the range: `in` = false
// makes its range empty).  Decodes the strip base over rings m+1.. in (h, l) form, sets up the
// lane's mask-row addresses, then sweeps the strip.
template <int F, int G, int W, int C>
__device__ __forceinline__ void es_strip(const HotPlan& P, const uint32_t* lds, uint64_t u,
                                         bool in, uint64_t o_lo, uint64_t o_hi, uint64_t idx_lo,
                                         uint64_t idx_hi) {
  const uint64_t M0l = P.M0_lo, M0h = P.M0_hi;
  const uint32_t q0 = P.q0, S = P.S;
  const uint32_t row_bytes = 8u * (W ? (uint32_t)W : P.W);
  cu32_t* mk = (cu32_t*)(P.masks);
  uint32_t hu = 0;
  uint64_t ll = 0, lh = 0;
  uint64_t rem = u;
  for (uint32_t j = 0; j < P.nb; ++j) {
    const uint32_t r = P.base_radix[j];
    uint32_t d;
    if (j + 1 == P.nb) {
      d = (uint32_t)rem;
    } else {
      const uint64_t qt = rem / r;
      d = (uint32_t)(rem - qt * r);
      rem = qt;
    }
    const uint32_t o = P.base_off[j] + d;
    const uint64_t tl = P.term_l[2 * o], th = P.term_l[2 * o + 1];
    const uint64_t nl = ll + tl;
    lh = lh + th + (nl < ll ? 1ull : 0ull);
    ll = nl;
    uint32_t nh = hu + P.term_h[o];
    if (es_ge128(lh, ll, M0h, M0l)) {
      const uint64_t bl = ll < M0l ? 1ull : 0ull;
      ll -= M0l;
      lh = lh - M0h - bl;
      nh += 1u;
    }
    if (nh >= q0) nh -= q0;
    hu = nh;
  }
  const uint64_t zl = M0l - ll;  // z = M0 - l_u: l_u + L >= M0 <=> L >= z
  const uint64_t zh = M0h - lh - (M0l < ll ? 1ull : 0ull);
  // byte address of the row of this lane's strip-base phase x in each mask table
  uint32_t A[F + 1];
#pragma unroll
  for (int f = 0; f < F; ++f) {
    cu32_t* rec = mk + kMaskInfo * f;
    const uint32_t m = rec[0];
    const uint64_t bm = (uint64_t)hu * rec[2] + es_mod128(ll, lh, m, rec[3], rec[4], rec[5]) +
                        rec[7];
    const uint32_t x = (uint32_t)(bm % m);
    A[f] = rec[1] + row_bytes * (uint32_t)((uint64_t)x * rec[6] % m);
  }
  A[F] = 0;
  const uint64_t obase = u * S;
  uint2* wq = es_wave_queue(P, lds);
  uint32_t sb = 0, se = 0;
  if (in) {
    sb = o_lo > obase ? (uint32_t)(o_lo - obase) : 0u;
    se = o_hi - obase >= S ? S : (uint32_t)(o_hi - obase);
  }
  const bool full = in && sb == 0 && se == S;
  if (__all(full))
    es_sweep<F, G, W, C, false>(P, lds, A, hu, zl, zh, obase, sb, se, wq, idx_lo, idx_hi);
  else
    es_sweep<F, G, W, C, true>(P, lds, A, hu, zl, zh, obase, sb, se, wq, idx_lo, idx_hi);
}

// One-word rows fit 80 VGPRs (6 waves per SIMD); wider rows keep 5 waves rather than spill
// their unrolled word loops.
template <int F, int G, int W, int C>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(W == 1 ? 6 : 5)))
es_sieve_kernel(const HotPlan P, uint64_t u_lo, uint64_t u_hi, uint64_t o_lo, uint64_t o_hi,
                uint64_t idx_lo, uint64_t idx_hi) {
  extern __shared__ uint32_t lds[];
  for (uint32_t z = threadIdx.x; z < P.lds_words; z += blockDim.x) lds[z] = P.lds_src[z];
  __syncthreads();
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  // Wave-uniform strip loop: a lane past u_hi replays strip u_lo with an empty range.
  for (uint64_t base = u_lo + (uint64_t)blockIdx.x * blockDim.x; base < u_hi; base += step) {
    const uint64_t my = base + threadIdx.x;
    const bool in = my < u_hi;
    es_strip<F, G, W, C>(P, lds, in ? my : u_lo, in, o_lo, o_hi, idx_lo, idx_hi);
  }
}

// Grid-mapped batch: one launch over all jobs of a shape; workgroup b serves strips
// [first_u + (b - wg_start) * blockDim.x, ...) of the job whose workgroup range holds b (found by
// binary search over the jobs' first workgroups).  The hardware dispatcher hands workgroups to
// CUs as they free up, so uneven jobs balance at workgroup granularity, as in the single-plan
// kernel.
struct GridJob {
  uint64_t wg_start;           // first workgroup of the job in the launch
  uint64_t u_lo, u_hi, o_lo, o_hi, idx_lo, idx_hi;
  uint32_t plan, pad;
};

template <int F, int G, int W, int C>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(W == 1 ? 6 : 5)))
es_grid_batch_kernel(const HotPlan* __restrict__ plans, const GridJob* __restrict__ jobs,
                     uint32_t n_jobs) {
  extern __shared__ uint32_t lds[];
  const uint64_t b = blockIdx.x;
  uint32_t lo = 0, hi = n_jobs;  // last job with wg_start <= b
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].wg_start <= b) lo = mid; else hi = mid;
  }
  const GridJob J = jobs[lo];
  const HotPlan P = plans[J.plan];
  for (uint32_t z = threadIdx.x; z < P.lds_words; z += blockDim.x) lds[z] = P.lds_src[z];
  __syncthreads();
  const uint64_t my = J.u_lo + (b - J.wg_start) * blockDim.x + threadIdx.x;
  const bool in = my < J.u_hi;
  es_strip<F, G, W, C>(P, lds, in ? my : J.u_lo, in, J.o_lo, J.o_hi, J.idx_lo, J.idx_hi);
}

// ============================================================================ row-lane layout
//
// Lanes over strip rows.  The strip rows sg in [0, S) (rings 1..m) are grouped 64 to a wave,
// chosen at engine creation so that within each half-wave of 32 lanes their rows of every mask
// table fall in distinct bank pairs as far as possible (a ds_read_b64 serves 32 lanes a cycle
// when their 8-byte rows sit in 32 distinct bank pairs).  A wave then walks V consecutive
// steps v of the sub rings m+1..m+n under one superstrip base w of the remaining rings: the
// step's values are one wave-uniform row of a step table (scalar loads), the superstrip's are
// decoded once per unit of V steps.  Outer index o = sg + S * (v + V * w).
//
// Table layout.  Mask table f keeps phase x at row x * lambda mod m, lambda = 32 * (-M)^-1
// mod m: subtracting M moves a phase 32 rows on, a whole round of the 64 banks (32 rows of 8
// bytes), so the wraps past M never move a lane to another bank.  (A modulus 2^D has no such
// lambda -- 32 is not invertible mod 2^D -- so a power-of-2 table, placed last, keeps
// lambda = (-M)^-1: one row per M subtracted, and its lanes may change bank on a wrap.)  A lane's row is
// Q_sg + W_w + 32 c_u + D_v + 32 c_v (strip row, superstrip base, the wrap of their sum at
// unit set-up, step, the wrap at the step) without reduction, so rows [0, 3m + 64) are
// stored; only Q_sg and c_u differ between the lanes of a wave, and c_u keeps the bank.
// (Reducing the lane's row mod m at unit set-up would halve that but moves the lanes whose
// sum passes m by -m rows, to another bank: 1.5x the conflict cycles at k=399, measured.)
// A ring 0 of two words (33..64 slots) keeps one such table per word (8-byte rows: the word
// of x and of x - M); the word-1 tables of all masks form a second block, word_bytes after
// the first, so a lane's word-1 row sits at the same bank offset from its word-0 row in every
// table and the grouping serves both words.  (Three and four words were built too: their
// instances need 96 VGPRs and spill, and they ran 0.64-0.98x the strip-lane layout on every
// frontier wheel measured, so they are not instantiated.)
constexpr int kRowsMaxMasks = 11;   // a step row: h, z (4 dwords), then one row offset per mask
constexpr int kStepDwords = 16;
constexpr int kRowsMaxWords = 2;
// Row-lane queue entry: {bits, v | w << kRowsWordShift | lane << kQueueIndexBits} (v < 1024:
// steps per unit; w: the ring-0 word)
constexpr uint32_t kRowsWordShift = 10;

struct RowPlan {
  uint64_t M0_lo, M0_hi;
  uint32_t q0, r0, S, G;           // ring-0 modulus and radix; strip rows; lane groups of 64
  uint32_t V, nb;                  // steps per unit; superstrip rings
  uint32_t lds_words, carry_base;  // LDS image size (words), carry table byte offset
  uint32_t stride;                 // dwords per strip row
  uint32_t last_even;              // 1: the last table's modulus is even (a power of 2): its
                                   // rows map x -> x (-M)^-1 (one row per M subtracted)
  uint32_t word_bytes;             // bytes from a mask's word-w table to its word-(w+1) table
  const uint32_t* lds_src;         // device copy of the LDS image
  const uint32_t* perm;            // [G * 64] strip row of each lane slot; bit 31: a padding
                                   // lane (it repeats a row of its half-wave and emits nothing)
  const uint32_t* strip;           // [S][stride] H, X, L (4 dwords), 0, 0, Q8_f (row bytes)
  const uint32_t* steps;           // [V][16] h, z = M0 - l (4 dwords), D8_f (row bytes)
  const uint32_t* base_radix;      // [nb]
  const uint32_t* base_off;        // [nb] first term of each superstrip ring
  const uint64_t* base_prefix;     // [nb] product of the radices of the rings before ring j
  const uint32_t* term_h;          // superstrip-ring terms, h part
  const uint64_t* term_l;          // superstrip-ring terms, l part (lo, hi pairs)
  const uint32_t* masks;           // [F][8] m, table byte base, M0 mod m, 2^32, 2^64, 2^96 mod m,
                                   //        lambda, t*M mod m
  const ColdPlan* cold;
};

__device__ __forceinline__ uint2* es_rows_queue(const RowPlan& P, const uint32_t* lds) {
  const uint32_t base = (P.lds_words * 4u + 15u) & ~15u;
  // the wave's index is uniform across its lanes: a scalar register holds the queue's address
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  return reinterpret_cast<uint2*>(reinterpret_cast<char*>(const_cast<uint32_t*>(lds)) + base +
                                  wave * kWaveQueueBytes);
}

// h and l of n's residue at step row `st` for a lane with superstrip-plus-strip-row sums
// reduced at unit set-up (Hp < q0, Lp < M0): b = (Hp + h_D) M0 + Lp + l_D, reduced to h < q0;
// c = 1 when M was subtracted.
__device__ __forceinline__ void es_rows_split(uint32_t q0, uint32_t Hp, uint64_t Lpl, uint64_t Lph,
                                              uint32_t hD, uint64_t zl, uint64_t zh,
                                              uint32_t& h, uint32_t& c, bool& kap) {
  kap = es_ge128(Lph, Lpl, zh, zl);          // Lp + l_D >= M0
  h = Hp + hD + (kap ? 1u : 0u);             // < 2 q0
  c = h >= q0 ? 1u : 0u;
  h = c ? h - q0 : h;
}

// Exact re-test of queued entries [first, first + n) (one per lane), called by the whole wave;
// an entry is {bits, v | w << 10 | lane << 26}: the entry's lane supplies its sums, strip row
// and correction through cross-lane permutes, its step row is read by vector loads.
template <int C, int W>
__device__ __forceinline__ void es_rows_drain(const RowPlan& P, const uint2* q, uint32_t first,
                                              uint32_t n, uint32_t Hp, uint64_t Lpl, uint64_t Lph,
                                              uint32_t sg, uint32_t X, uint64_t obase,
                                              uint64_t idx_lo, uint64_t idx_hi) {
  __builtin_amdgcn_wave_barrier();
  const uint32_t lane = threadIdx.x & 63u;
  const uint2 ent = lane < n ? q[first + lane] : make_uint2(0u, 0u);
  const uint32_t src = ent.y >> kQueueIndexBits;
  const uint64_t sLl = (uint64_t)es_from_lane(src, (uint32_t)Lpl) |
                       ((uint64_t)es_from_lane(src, (uint32_t)(Lpl >> 32)) << 32);
  const uint64_t sLh = (uint64_t)es_from_lane(src, (uint32_t)Lph) |
                       ((uint64_t)es_from_lane(src, (uint32_t)(Lph >> 32)) << 32);
  const uint32_t sH = es_from_lane(src, Hp), ssg = es_from_lane(src, sg), sX = es_from_lane(src, X);
  if (lane < n) {
    const uint32_t v = ent.y & ((1u << kRowsWordShift) - 1u);
    const uint32_t w = W == 1 ? 0u : (ent.y >> kRowsWordShift) & (kRowsMaxWords - 1u);
    const uint32_t* st = P.steps + (size_t)v * kStepDwords;
    const uint64_t zl = (uint64_t)st[1] | ((uint64_t)st[2] << 32);
    const uint64_t zh = (uint64_t)st[3] | ((uint64_t)st[4] << 32);
    uint32_t h, c;
    bool kap;
    es_rows_split(P.q0, sH, sLl, sLh, st[0], zl, zh, h, c, kap);
    const u128 Lp = es::mk128(sLl, sLh), z = es::mk128(zl, zh);
    const u128 M0 = es::mk128(P.M0_lo, P.M0_hi);
    const u128 l = kap ? Lp - z : Lp + (M0 - z);   // Lp + l_D (mod M0)
    const uint64_t o = obase + (uint64_t)v * P.S + ssg;
    const uint64_t ob = (C & 2) ? o * P.r0 + (uint64_t)(int64_t)(int32_t)sX : o;
    es_exact_test_hl<C>(P.cold, ent.x, w, h, l, ob, idx_lo, idx_hi);
  }
  __builtin_amdgcn_wave_barrier();
}

// One unit: lane group g of the strip rows, superstrip w, all V steps.  Wave-uniform.  W: words
// per ring-0 row (W >= 2 reads the byte table of j(h), C == 1).
template <int F, int C, bool CHECK, int W>
__device__ __forceinline__ void es_rows_sweep(const RowPlan& P, const uint32_t* lds,
                                              const uint32_t (&A)[F + 1], uint32_t Hp,
                                              uint64_t Lpl, uint64_t Lph, uint32_t sg,
                                              uint32_t X, bool valid, uint64_t obase,
                                              uint64_t o_lo, uint64_t o_hi, uint64_t idx_lo,
                                              uint64_t idx_hi) {
  const uint32_t q0 = P.q0, V = P.V, S = P.S;
  const char* lb = reinterpret_cast<const char*>(lds);
  // the carry table opens the LDS image (its address is h itself)
  const uint8_t* carry_tab = reinterpret_cast<const uint8_t*>(lds);
  const uint32_t* carry_w = lds;
  uint2* wq = es_rows_queue(P, lds);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t vmask = __ballot(valid);
  static_assert(W >= 1 && W <= kRowsMaxWords && (W == 1 || C == 1),
                "wider rows read the byte carry table and have no lead block");
  const uint32_t wb = W > 1 ? P.word_bytes : 0u;
  uint32_t cnt = 0;
  // Each lane holds its latest candidate word (bits, step and word) in registers; `pend` marks
  // the lanes holding one (wave-uniform).  Candidate words are rare per lane (about one in 64
  // steps at k=399), so the wave appends all pending entries to its LDS queue at once, when a
  // lane finds a second one, instead of one LDS write per step that has any.
  uint32_t pbits = 0, py = 0;
  uint64_t pend = 0;
  auto flush = [&]() {
    if ((pend >> lane) & 1ull) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(
          (uint32_t)(pend >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)pend, 0u));
      wq[cnt + below] = make_uint2(pbits, py);
    }
    cnt += (uint32_t)__popcll(pend);
    pend = 0;
    if (cnt >= 64u) {  // at most 127 queued: a drain follows every flush that passes 64
      cnt -= 64u;
      es_rows_drain<C, W>(P, wq, cnt, 64u, Hp, Lpl, Lph, sg, X, obase, idx_lo, idx_hi);
    }
  };
  for (uint32_t v = 0; v < V; ++v) {
    cu32_t* st = (cu32_t*)(P.steps) + (size_t)v * kStepDwords;
    const uint64_t zl = (uint64_t)st[1] | ((uint64_t)st[2] << 32);
    const uint64_t zh = (uint64_t)st[3] | ((uint64_t)st[4] << 32);
    uint32_t h, c;
    bool kap;
    es_rows_split(q0, Hp, Lpl, Lph, st[0], zl, zh, h, c, kap);
    const uint32_t wrap = c ? 256u : 0u;      // 32 rows of 8 bytes when M was subtracted
    const uint32_t wrap_last = c ? (P.last_even ? 8u : 256u) : 0u;
    uint32_t bx[W], by[W];
#pragma unroll
    for (int w = 0; w < W; ++w) bx[w] = by[w] = 0xffffffffu;
#pragma unroll
    for (int f = 0; f < F; ++f) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint2 x = *reinterpret_cast<const uint2*>(
            lb + A[f] + st[5 + f] + (f == F - 1 ? wrap_last : wrap) + (uint32_t)w * wb);
        bx[w] &= x.x;
        by[w] &= x.y;
      }
    }
    // bits i >= j(h) of the ring-0 row carry past M
    uint32_t carry[W];
    if constexpr (W == 1) {
      if (C & 1) {  // byte table of j(h) <= 32: one 64-bit shift
        carry[0] = (uint32_t)(0xffffffffull << carry_tab[h]);
      } else {
        carry[0] = carry_w[h];
      }
    } else {        // byte table of j(h) <= 64: word w carries from bit j - 32 w on
      const int j = (int)carry_tab[h];
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const int sh = j - 32 * w;
        carry[w] = sh <= 0 ? 0xffffffffu : sh >= 32 ? 0u : 0xffffffffu << sh;
      }
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint32_t bits = (carry[w] & by[w]) | (~carry[w] & bx[w]);
      if (CHECK) {
        const uint64_t o = obase + (uint64_t)v * S + sg;
        if (o < o_lo || o >= o_hi) bits = 0;
      }
      // padding lanes (a repeated row) are left out through the wave's mask of valid lanes
      const uint64_t any = __ballot(bits != 0u) & vmask;
      if (any) {
        if (any & pend) flush();   // a lane's slot is taken: every pending entry to the queue
        if (bits && valid) {
          pbits = bits;
          py = v | ((uint32_t)w << kRowsWordShift) | (lane << kQueueIndexBits);
        }
        pend |= any;
      }
    }
  }
  if (pend) flush();
  if (cnt) es_rows_drain<C, W>(P, wq, 0, cnt, Hp, Lpl, Lph, sg, X, obase, idx_lo, idx_hi);
}

// Unit set-up: decode superstrip w, its table rows (block constant included), the lane's
// strip row and sums; then sweep the V steps.
template <int F, int C, int W>
__device__ __forceinline__ void es_rows_unit(const RowPlan& P, const uint32_t* lds, uint32_t g,
                                             uint64_t w, uint64_t o_lo, uint64_t o_hi,
                                             uint64_t idx_lo, uint64_t idx_hi) {
  const uint64_t M0l = P.M0_lo, M0h = P.M0_hi;
  const uint32_t q0 = P.q0;
  // Superstrip base in parallel: lane j < nb takes the digit of ring j (w / prefix_j mod r_j)
  // and its term in (h, l) form; a butterfly over the lanes then sums the terms mod M, so every
  // lane ends with (hu, l) of the base.
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nb = P.nb;
  uint32_t hu = 0;
  uint64_t ll = 0, lh = 0;
  if (lane < nb) {
    const uint64_t qv = w / P.base_prefix[lane];
    const uint32_t d = lane + 1 == nb ? (uint32_t)qv : (uint32_t)(qv % P.base_radix[lane]);
    const uint32_t o = P.base_off[lane] + d;
    hu = P.term_h[o];
    ll = P.term_l[2 * o];
    lh = P.term_l[2 * o + 1];
  }
  for (uint32_t sh = 1; sh < nb; sh <<= 1) {
    const uint32_t src = lane ^ sh;
    const uint32_t oh = es_from_lane(src, hu);
    const uint64_t ol = (uint64_t)es_from_lane(src, (uint32_t)ll) |
                        ((uint64_t)es_from_lane(src, (uint32_t)(ll >> 32)) << 32);
    const uint64_t olh = (uint64_t)es_from_lane(src, (uint32_t)lh) |
                         ((uint64_t)es_from_lane(src, (uint32_t)(lh >> 32)) << 32);
    const uint64_t nl = ll + ol;
    lh = lh + olh + (nl < ll ? 1ull : 0ull);
    ll = nl;
    uint32_t nh = hu + oh;
    if (es_ge128(lh, ll, M0h, M0l)) {
      const uint64_t bl = ll < M0l ? 1ull : 0ull;
      ll -= M0l;
      lh = lh - M0h - bl;
      nh += 1u;
    }
    if (nh >= q0) nh -= q0;
    if (nh >= q0) nh -= q0;
    hu = nh;
  }
  // lane 0's block holds the total: make it wave-uniform
  hu = (uint32_t)__builtin_amdgcn_readlane((int)hu, 0);
  ll = es_readlane64(ll, 0);
  lh = es_readlane64(lh, 0);
  // row of the base's phase (block constant included) in table f, computed by lane f
  cu32_t* mk = (cu32_t*)(P.masks);
  uint32_t rowf = 0;
  if (lane < (uint32_t)F) {
    const uint32_t* rec = P.masks + kMaskInfo * lane;
    const uint32_t m = rec[0];
    const uint64_t bm = (uint64_t)hu * rec[2] + es_mod128(ll, lh, m, rec[3], rec[4], rec[5]) +
                        rec[7];
    rowf = (uint32_t)((uint64_t)(uint32_t)(bm % m) * rec[6] % m);
  }
  const uint32_t pe = P.perm[(size_t)g * 64u + lane];
  const uint32_t sg = pe & 0x7fffffffu;
  const bool valid = (pe >> 31) == 0;
  cu32_t* srow = (cu32_t*)P.strip + (size_t)sg * P.stride;
  const uint64_t Ll = (uint64_t)srow[2] | ((uint64_t)srow[3] << 32);
  const uint64_t Lh = (uint64_t)srow[4] | ((uint64_t)srow[5] << 32);
  // Lp = L_sg + l_w (mod M0), Hp = H_sg + h_w (+1 on the carry), then reduced below q0: a
  // lane whose strip row plus superstrip passes M starts 32 rows on in every table
  uint64_t Lpl = Ll + ll;
  uint64_t Lph = Lh + lh + (Lpl < Ll ? 1ull : 0ull);
  uint32_t Hp = srow[0] + hu;
  if (es_ge128(Lph, Lpl, M0h, M0l)) {
    const uint64_t bl = Lpl < M0l ? 1ull : 0ull;
    Lpl -= M0l;
    Lph = Lph - M0h - bl;
    Hp += 1u;
  }
  // the lane's row in table f: strip row + superstrip row (+ 32 rows, or one for an even last
  // modulus, when M was subtracted), left unreduced: a reduction mod m would move the lanes
  // whose sum passes m to another bank (m is odd), undoing the grouping of the strip rows
  const uint32_t lane_wrap = Hp >= q0 ? 256u : 0u;
  const uint32_t lane_wrap_last = Hp >= q0 ? (P.last_even ? 8u : 256u) : 0u;
  if (Hp >= q0) Hp -= q0;
  uint32_t A[F + 1];
#pragma unroll
  for (int f = 0; f < F; ++f)
    A[f] = mk[kMaskInfo * f + 1] + srow[8 + f] +
           8u * (uint32_t)__builtin_amdgcn_readlane((int)rowf, f) +
           (f == F - 1 ? lane_wrap_last : lane_wrap);
  A[F] = 0;
  const uint64_t span = (uint64_t)P.V * P.S;
  const uint64_t obase = w * span;
  const bool full = obase >= o_lo && obase + span <= o_hi;
  if (full)
    es_rows_sweep<F, C, false, W>(P, lds, A, Hp, Lpl, Lph, sg, srow[1], valid, obase, o_lo,
                                  o_hi, idx_lo, idx_hi);
  else
    es_rows_sweep<F, C, true, W>(P, lds, A, Hp, Lpl, Lph, sg, srow[1], valid, obase, o_lo,
                                 o_hi, idx_lo, idx_hi);
}

// Units g + G * (w - w_lo) for w in [w_lo, w_hi), one per wave: the grid covers them all (the
// host splits a range whose units exceed one grid).  WPE: waves per SIMD the registers are
// sized for.  One-word rows come at 6 (78 VGPRs, no spills) and at 8 (64 VGPRs; the exact
// re-test of the drain spills 5-24 of them, the step loop none): the engine takes the 8-wave
// instance only where it keeps more waves resident, i.e. where two 1024-thread workgroups'
// tables fit the CU's LDS (k=399: 2.07 against 2.18 ms per 2^36 launch; where the occupancy
// does not change the spills cost up to 4 %).
template <int F, int C, int W, int WPE>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(WPE)))
es_rows_kernel(const RowPlan* __restrict__ Pp, uint64_t w_lo, uint64_t w_hi, uint64_t o_lo,
               uint64_t o_hi, uint64_t idx_lo, uint64_t idx_hi) {
  const RowPlan& P = *Pp;
  extern __shared__ uint32_t lds[];
  for (uint32_t z = threadIdx.x; z < P.lds_words; z += blockDim.x) lds[z] = P.lds_src[z];
  __syncthreads();
  const uint32_t wpg = blockDim.x >> 6;
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t units = (uint64_t)P.G * (w_hi - w_lo);
  const uint64_t q = (uint64_t)blockIdx.x * wpg + wave;
  if (q < units) {
    const uint64_t wq = q / P.G;
    es_rows_unit<F, C, W>(P, lds, (uint32_t)(q - wq * P.G), w_lo + wq, o_lo, o_hi, idx_lo,
                          idx_hi);
  }
}

// Grid-mapped batch of row-lane jobs: workgroup b serves units (b - wg_start) * waves + wave
// of the job whose workgroup range holds b.  GridJob's u_lo/u_hi hold the job's superstrips.
template <int F, int C, int W, int WPE>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(WPE)))
es_rows_batch_kernel(const RowPlan* __restrict__ plans, const GridJob* __restrict__ jobs,
                     uint32_t n_jobs) {
  extern __shared__ uint32_t lds[];
  const uint64_t b = blockIdx.x;
  uint32_t lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (jobs[mid].wg_start <= b) lo = mid; else hi = mid;
  }
  const GridJob J = jobs[lo];
  const RowPlan& P = plans[J.plan];
  for (uint32_t z = threadIdx.x; z < P.lds_words; z += blockDim.x) lds[z] = P.lds_src[z];
  __syncthreads();
  const uint32_t wpg = blockDim.x >> 6;
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t q = (b - J.wg_start) * wpg + wave;
  if (q < (uint64_t)P.G * (J.u_hi - J.u_lo)) {
    const uint64_t wq = q / P.G;
    es_rows_unit<F, C, W>(P, lds, (uint32_t)(q - wq * P.G), J.u_lo + wq, J.o_lo, J.o_hi,
                          J.idx_lo, J.idx_hi);
  }
}

typedef void (*RowsFn)(const RowPlan*, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t);
typedef void (*RowsBatchFn)(const RowPlan*, const GridJob*, uint32_t);

// Row-lane instances: W = 1 with each carry form and lead-block form (C = 0..3), sized for 6
// or (wave8) 8 waves per SIMD; W = 2 with the byte carry table (C = 1), 5 waves.
template <int F>
RowsFn rows_kernel_for(uint32_t C, uint32_t W, bool wave8) {
  if (F < 1 || F > kRowsMaxMasks) return nullptr;
  if (W != 1) return W == 2 && C == 1 && !wave8 ? es_rows_kernel<F, 1, 2, 5> : nullptr;
  switch (C) {
    case 0: return wave8 ? es_rows_kernel<F, 0, 1, 8> : es_rows_kernel<F, 0, 1, 6>;
    case 1: return wave8 ? es_rows_kernel<F, 1, 1, 8> : es_rows_kernel<F, 1, 1, 6>;
    case 2: return wave8 ? es_rows_kernel<F, 2, 1, 8> : es_rows_kernel<F, 2, 1, 6>;
    case 3: return wave8 ? es_rows_kernel<F, 3, 1, 8> : es_rows_kernel<F, 3, 1, 6>;
    default: return nullptr;
  }
}

template <int F>
RowsBatchFn rows_batch_for(uint32_t C, uint32_t W, bool wave8) {
  if (F < 1 || F > kRowsMaxMasks) return nullptr;
  if (W != 1) return W == 2 && C == 1 && !wave8 ? es_rows_batch_kernel<F, 1, 2, 5> : nullptr;
  switch (C) {
    case 0: return wave8 ? es_rows_batch_kernel<F, 0, 1, 8> : es_rows_batch_kernel<F, 0, 1, 6>;
    case 1: return wave8 ? es_rows_batch_kernel<F, 1, 1, 8> : es_rows_batch_kernel<F, 1, 1, 6>;
    case 2: return wave8 ? es_rows_batch_kernel<F, 2, 1, 8> : es_rows_batch_kernel<F, 2, 1, 6>;
    case 3: return wave8 ? es_rows_batch_kernel<F, 3, 1, 8> : es_rows_batch_kernel<F, 3, 1, 6>;
    default: return nullptr;
  }
}

typedef void (*KernelFn)(HotPlan, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t);
typedef void (*GridBatchFn)(const HotPlan*, const GridJob*, uint32_t);

// Instances: every F <= 12 with all masks in LDS for W = 1, 2, 3, 4 and runtime W; for W = 1
// also G = 2 or 4 of the masks in global memory, and (G = 0) the byte carry table (C & 1)
// and the lead block's idx form (C & 2).
template <int F>
KernelFn kernel_for(uint32_t W, uint32_t G, uint32_t C) {
  if (C) {
    if (W != 1 || G != 0) return nullptr;
    if (C == 1) return es_sieve_kernel<F, 0, 1, 1>;
    if (C == 2) return es_sieve_kernel<F, 0, 1, 2>;
    return C == 3 ? es_sieve_kernel<F, 0, 1, 3> : nullptr;
  }
  if (W == 1) {
    if (G == 2 && F >= 2) return es_sieve_kernel<F, (F >= 2 ? 2 : 0), 1, 0>;
    if (G == 4 && F >= 4) return es_sieve_kernel<F, (F >= 4 ? 4 : 0), 1, 0>;
    return G == 0 ? es_sieve_kernel<F, 0, 1, 0> : nullptr;
  }
  if (G != 0) return nullptr;
  if (W == 2) return es_sieve_kernel<F, 0, 2, 0>;
  if (W == 3) return es_sieve_kernel<F, 0, 3, 0>;
  if (W == 4) return es_sieve_kernel<F, 0, 4, 0>;
  return es_sieve_kernel<F, 0, 0, 0>;
}

template <int F>
GridBatchFn grid_batch_for(uint32_t W, uint32_t G, uint32_t C) {
  if (G != 0) return nullptr;
  if (C) {
    if (W != 1) return nullptr;
    if (C == 1) return es_grid_batch_kernel<F, 0, 1, 1>;
    if (C == 2) return es_grid_batch_kernel<F, 0, 1, 2>;
    return C == 3 ? es_grid_batch_kernel<F, 0, 1, 3> : nullptr;
  }
  if (W == 1) return es_grid_batch_kernel<F, 0, 1, 0>;
  if (W == 2) return es_grid_batch_kernel<F, 0, 2, 0>;
  if (W == 3) return es_grid_batch_kernel<F, 0, 3, 0>;
  if (W == 4) return es_grid_batch_kernel<F, 0, 4, 0>;
  return es_grid_batch_kernel<F, 0, 0, 0>;
}


// Instance selectors of the four units (F outside a unit's range: nullptr).
KernelFn pick_kernel_p0(int F, uint32_t W, uint32_t G, uint32_t C);
KernelFn pick_kernel_p1(int F, uint32_t W, uint32_t G, uint32_t C);
KernelFn pick_kernel_p2(int F, uint32_t W, uint32_t G, uint32_t C);
KernelFn pick_kernel_p3(int F, uint32_t W, uint32_t G, uint32_t C);
GridBatchFn pick_grid_batch_p0(int F, uint32_t W, uint32_t G, uint32_t C);
GridBatchFn pick_grid_batch_p1(int F, uint32_t W, uint32_t G, uint32_t C);
GridBatchFn pick_grid_batch_p2(int F, uint32_t W, uint32_t G, uint32_t C);
GridBatchFn pick_grid_batch_p3(int F, uint32_t W, uint32_t G, uint32_t C);
RowsFn pick_rows_p0(int F, uint32_t C, uint32_t W, bool wave8);
RowsFn pick_rows_p1(int F, uint32_t C, uint32_t W, bool wave8);
RowsFn pick_rows_p2(int F, uint32_t C, uint32_t W, bool wave8);
RowsFn pick_rows_p3(int F, uint32_t C, uint32_t W, bool wave8);
RowsBatchFn pick_rows_batch_p0(int F, uint32_t C, uint32_t W, bool wave8);
RowsBatchFn pick_rows_batch_p1(int F, uint32_t C, uint32_t W, bool wave8);
RowsBatchFn pick_rows_batch_p2(int F, uint32_t C, uint32_t W, bool wave8);
RowsBatchFn pick_rows_batch_p3(int F, uint32_t C, uint32_t W, bool wave8);

// Mask counts served by each instance unit.
constexpr int kPartFirstF[5] = {0, 4, 7, 10, 13};

}  // namespace es_dev
