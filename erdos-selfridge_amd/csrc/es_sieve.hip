// es_sieve.hip -- MI355X (gfx950) CRT-wheel sieve engine: device kernels and the C ABI of
// include/es_sieve.h.
//
// Enumeration.  A wheel of rings j = 0..w-1 (moduli q_j, admissible residues, CRT terms
// term_j[d] = (a_d * basis_j) mod M) maps a position idx = d_0 + r_0*(d_1 + r_1*(d_2 + ...))
// of block t to n = t*M + s, s = (sum_j term_j[d_j]) mod M  (sieve/gpu_sieve_v9.py:385-390).
// The outer index o = idx / r0 fixes d_1.. ; ring 0 is swept inside the thread.
//
// Mixed representation.  Ring-0 terms are exact multiples of M0 = M / q0:
// term_0[i] = e_i * M0 with e_i = (a_i * (M0^-1 mod q0)) mod q0.  Every other residue b in
// [0, M) is held as (h, l) = (b div M0, b mod M0), h < q0.  Then b + term_0[i] wraps past M
// iff h + e_i >= q0, so the ring-0 wrap flag of every position of an outer index follows
// from the small integer h alone (a split table lookup), never from a 128-bit compare.
//
// Strips.  Rings 1..m form a strip of S = r_1*...*r_m outer indices that every lane of a
// wave walks in the same order, so the strip's partial sums (H_sigma, L_sigma) and their
// residues mod each mask modulus are wave-uniform table rows (scalar loads).  Rings m+1..
// are the lane's strip base u, decoded once per strip.  Per outer index a lane does one
// 128-bit compare (carry of l), a few 32-bit adds, and for each word-mask prefilter one LDS
// row load: row x = (b + tM) mod m_f holds two r0-bit masks (ring-0 wrap clear / set)
// of the positions whose n passes the prefilter, combined through the wrap mask.
//
// Exactness.  Prefilters are necessary conditions (low base-p digits of a filter prime),
// so they only discard positions the full predicate rejects.  Every surviving position is
// then re-tested with the full admissibility tables of all slow filters from its exact
// 128-bit residue s, which is the reference predicate (sieve/gpu_sieve_v9.py:304-322).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/es_sieve.h"

typedef unsigned __int128 u128;

namespace {

constexpr int kMaxMasks = 12;

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define ES_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t err_ = (call);                                                              \
    if (err_ != hipSuccess)                                                                \
      return fail(ES_ERR_DEVICE, std::string(#call) + " failed: " + hipGetErrorString(err_)); \
  } while (0)

// Everything a sieve launch reads.  Passed by value (kernel arguments are scalar loads).
struct DevPlan {
  uint64_t M0_lo, M0_hi;       // M0 = M / q0
  uint32_t q0, r0, W, S;       // ring-0 modulus, radix, words per mask row, strip length
  uint32_t nb, n_slow;         // base rings (m+1..w-1), exact filters
  uint32_t lds_words;          // LDS image size (mask rows, then the split table)
  uint32_t split_off;          // word offset of the split table in LDS
  const uint32_t* lds_src;     // device copy of the LDS image
  const uint32_t* base_radix;  // [nb]
  const uint32_t* base_off;    // [nb] first term of each base ring
  const uint32_t* term_h;      // base-ring terms, h part
  const uint64_t* term_l;      // base-ring terms, l part (lo, hi pairs)
  const uint32_t* strip_h;     // [S] H_sigma
  const uint64_t* strip_l;     // [2S] L_sigma (lo, hi)
  const uint32_t* strip_q;     // [S][F][2] Q_sigma_f and (Q_sigma_f - M) mod m_f
  const uint32_t* mask_mod;    // [F]
  const uint32_t* mask_row;    // [F] LDS word offset of each mask's rows
  const uint32_t* mask_pow;    // [F][4] M0 mod m, 2^32, 2^64, 2^96 mod m
  const uint32_t* mask_tc;     // [F] t*M mod m
  const uint32_t* e0;          // [r0] ring-0 e values, ascending
  const uint32_t* inner_ids;   // [r0] original ring-0 digit of each sorted slot
  const uint32_t* slow_q;      // [n_slow]
  const uint32_t* slow_off;    // [n_slow] word offset into slow_bits
  const uint32_t* slow_bits;   // admissibility bitmaps
  const uint32_t* slow_pow;    // [n_slow][3] 2^32, 2^64, 2^96 mod q
  const uint32_t* slow_tc;     // [n_slow] t*M mod q
};

__device__ __forceinline__ bool ge128(uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) {
  return (ah > bh) || (ah == bh && al >= bl);
}

// x mod q for a 128-bit x given 2^32k mod q (q < 2^24 keeps the sum below 2^58).
__device__ __forceinline__ uint32_t mod128(uint64_t lo, uint64_t hi, uint32_t q, uint32_t p1,
                                           uint32_t p2, uint32_t p3) {
  uint64_t x = (uint64_t)(uint32_t)lo + (uint64_t)(uint32_t)(lo >> 32) * p1 +
               (uint64_t)(uint32_t)hi * p2 + (uint64_t)(uint32_t)(hi >> 32) * p3;
  return (uint32_t)(x % q);
}

// Exact re-test of the candidates in `bits` (sorted ring-0 slots 32*w + b) of outer index o.
__device__ __forceinline__ void exact_tail(const DevPlan& P, uint32_t bits, uint32_t w,
                                           uint32_t h, uint64_t l_lo, uint64_t l_hi,
                                           uint64_t o, uint64_t idx_lo, uint64_t idx_hi,
                                           unsigned long long* out,
                                           unsigned long long* out_n, uint64_t out_cap) {
  const u128 M0 = ((u128)P.M0_hi << 64) | P.M0_lo;
  const u128 l = ((u128)l_hi << 64) | l_lo;
  while (bits) {
    const uint32_t i = 32u * w + (uint32_t)__builtin_ctz(bits);
    bits &= bits - 1u;
    const uint64_t idx = o * P.r0 + P.inner_ids[i];
    if (idx < idx_lo || idx >= idx_hi) continue;
    uint32_t hh = h + P.e0[i];
    if (hh >= P.q0) hh -= P.q0;
    const u128 s = (u128)hh * M0 + l;
    const uint64_t s_lo = (uint64_t)s, s_hi = (uint64_t)(s >> 64);
    bool ok = true;
    for (uint32_t j = 0; j < P.n_slow; ++j) {
      const uint32_t q = P.slow_q[j];
      uint32_t r = mod128(s_lo, s_hi, q, P.slow_pow[3 * j], P.slow_pow[3 * j + 1],
                          P.slow_pow[3 * j + 2]) + P.slow_tc[j];
      if (r >= q) r -= q;
      if (!((P.slow_bits[P.slow_off[j] + (r >> 5)] >> (r & 31u)) & 1u)) {
        ok = false;
        break;
      }
    }
    if (ok) {
      const unsigned long long slot = atomicAdd(out_n, 1ull);
      if (slot < out_cap) out[slot] = idx;
    }
  }
}

// One strip sweep.  CHECK = the lane's strip is only partly inside [o_lo, o_hi).
template <int F, bool CHECK>
__device__ __forceinline__ void sweep_strip(const DevPlan& P, const uint32_t* lds,
                                            const uint32_t (&mm)[F + 1],
                                            const uint32_t (&mrow)[F + 1],
                                            uint32_t hu, uint64_t zl, uint64_t zh,
                                            const uint32_t (&bqu)[F + 1],
                                            uint64_t obase, uint32_t sb, uint32_t se,
                                            uint64_t idx_lo, uint64_t idx_hi,
                                            unsigned long long* out, unsigned long long* out_n,
                                            uint64_t out_cap) {
  const uint32_t q0 = P.q0, W = P.W, S = P.S, r0 = P.r0;
  const uint32_t* split = lds + P.split_off;
  for (uint32_t sg = 0; sg < S; ++sg) {
    const uint32_t Hs = P.strip_h[sg];
    const uint64_t Ll = P.strip_l[2 * sg], Lh = P.strip_l[2 * sg + 1];
    const bool kap = ge128(Lh, Ll, zh, zl);           // l_u + L_sigma >= M0
    uint32_t h = hu + Hs + (kap ? 1u : 0u);
    const bool c1 = h >= q0;                          // b_u + B_sigma >= M
    if (c1) h -= q0;
    uint32_t x[F + 1];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const uint32_t qa = P.strip_q[(sg * F + f) * 2], qb = P.strip_q[(sg * F + f) * 2 + 1];
      uint32_t v = bqu[f] + (c1 ? qb : qa);
      x[f] = min(v, v - mm[f]);
    }
    const uint32_t sp = split[h];                     // slots >= sp wrap past M
    bool live = true;
    if (CHECK) live = (sg >= sb) && (sg < se);
    for (uint32_t w = 0; w < W; ++w) {
      const int rel = (int)sp - (int)(32u * w);
      const uint32_t carry = rel <= 0 ? 0xffffffffu : (rel >= 32 ? 0u : (0xffffffffu << rel));
      uint32_t bits;
      if (F == 0) {
        const uint32_t n = r0 - 32u * w;
        bits = n >= 32u ? 0xffffffffu : ((1u << n) - 1u);
      } else {
        bits = 0xffffffffu;
      }
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const uint2 row = *reinterpret_cast<const uint2*>(lds + mrow[f] + 2u * (x[f] * W + w));
        bits &= (carry & row.y) | (~carry & row.x);
      }
      if (CHECK && !live) bits = 0;
      if (bits) {
        // exact residue of the outer index: l = l_u + L_sigma - kap*M0, with l_u = M0 - z
        const u128 M0 = ((u128)P.M0_hi << 64) | P.M0_lo;
        const u128 z = ((u128)zh << 64) | zl;
        const u128 L = ((u128)Lh << 64) | Ll;
        u128 l = (M0 - z) + L;
        if (kap) l -= M0;
        exact_tail(P, bits, w, h, (uint64_t)l, (uint64_t)(l >> 64), obase + sg, idx_lo, idx_hi,
                   out, out_n, out_cap);
      }
    }
  }
}

template <int F>
__global__ void __launch_bounds__(1024)
es_sieve_kernel(DevPlan P, uint64_t u_lo, uint64_t u_hi, uint64_t o_lo, uint64_t o_hi,
                uint64_t idx_lo, uint64_t idx_hi, unsigned long long* out,
                unsigned long long* out_n, uint64_t out_cap) {
  extern __shared__ uint32_t lds[];
  for (uint32_t z = threadIdx.x; z < P.lds_words; z += blockDim.x) lds[z] = P.lds_src[z];
  __syncthreads();

  uint32_t mm[F + 1], mrow[F + 1];
#pragma unroll
  for (int f = 0; f < F; ++f) {
    mm[f] = P.mask_mod[f];
    mrow[f] = P.mask_row[f];
  }
  mm[F] = 0;
  mrow[F] = 0;
  const uint64_t M0l = P.M0_lo, M0h = P.M0_hi;
  const uint32_t q0 = P.q0, S = P.S;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;

  // Wave-uniform strip loop: a lane past u_hi replays strip u_lo with an empty range.
  for (uint64_t base = u_lo + (uint64_t)blockIdx.x * blockDim.x; base < u_hi; base += step) {
    const uint64_t my = base + threadIdx.x;
    const bool in = my < u_hi;
    const uint64_t u = in ? my : u_lo;
    // decode the strip base over rings m+1..w-1 and sum their terms in (h, l) form
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
      if (ge128(lh, ll, M0h, M0l)) {
        const uint64_t bl = ll < M0l ? 1ull : 0ull;
        ll -= M0l;
        lh = lh - M0h - bl;
        nh += 1u;
      }
      if (nh >= q0) nh -= q0;
      hu = nh;
    }
    // z = M0 - l_u  (l_u + L >= M0  <=>  L >= z)
    const uint64_t zl = M0l - ll;
    const uint64_t zh = M0h - lh - (M0l < ll ? 1ull : 0ull);
    uint32_t bqu[F + 1];
#pragma unroll
    for (int f = 0; f < F; ++f) {
      const uint32_t m = mm[f];
      const uint32_t* pw = P.mask_pow + 4 * f;
      const uint64_t hm = (uint64_t)hu * pw[0];
      const uint32_t lm = mod128(ll, lh, m, pw[1], pw[2], pw[3]);
      bqu[f] = (uint32_t)((hm + lm + P.mask_tc[f]) % m);
    }
    bqu[F] = 0;
    const uint64_t obase = u * S;
    uint32_t sb = 0, se = 0;
    if (in) {
      sb = o_lo > obase ? (uint32_t)(o_lo - obase) : 0u;
      se = o_hi - obase >= S ? S : (uint32_t)(o_hi - obase);
    }
    const bool full = in && sb == 0 && se == S;
    if (__all(full)) {
      sweep_strip<F, false>(P, lds, mm, mrow, hu, zl, zh, bqu, obase, sb, se, idx_lo, idx_hi,
                            out, out_n, out_cap);
    } else {
      sweep_strip<F, true>(P, lds, mm, mrow, hu, zl, zh, bqu, obase, sb, se, idx_lo, idx_hi,
                           out, out_n, out_cap);
    }
  }
}

typedef void (*KernelFn)(DevPlan, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t, uint64_t,
                         unsigned long long*, unsigned long long*, uint64_t);

template <int F>
KernelFn kernel_for() {
  return es_sieve_kernel<F>;
}

KernelFn pick_kernel(int F) {
  switch (F) {
    case 0: return kernel_for<0>();
    case 1: return kernel_for<1>();
    case 2: return kernel_for<2>();
    case 3: return kernel_for<3>();
    case 4: return kernel_for<4>();
    case 5: return kernel_for<5>();
    case 6: return kernel_for<6>();
    case 7: return kernel_for<7>();
    case 8: return kernel_for<8>();
    case 9: return kernel_for<9>();
    case 10: return kernel_for<10>();
    case 11: return kernel_for<11>();
    case 12: return kernel_for<12>();
    default: return nullptr;
  }
}

// ------------------------------------------------------------------------------ host side

inline u128 mk128(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

uint32_t mod_u128(u128 x, uint32_t q) { return (uint32_t)(x % q); }

struct Blob {
  std::vector<uint8_t> bytes;
  size_t put(const void* p, size_t n) {
    size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + std::max<size_t>(n, 4), 0);
    if (n && p) std::memcpy(bytes.data() + off, p, n);
    return off;
  }
};

}  // namespace

struct es_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t k = 0, w = 0, r0 = 1, q0 = 1, W = 1, S = 1, F = 0, threads = 256;
  uint32_t n_filters = 0, n_slow = 0, lds_bytes = 0, max_grid = 1;
  u128 M = 1, M0 = 1, R = 1;
  uint64_t R_outer = 1, strips = 1, out_cap = 1u << 20, launch_outer = 1;
  std::vector<uint32_t> filt_q, slow_order, mask_mod, mask_src, radix;
  std::vector<u128> terms;          // all ring terms, ring-major (host-side idx -> s)
  std::vector<uint32_t> term_off;
  uint64_t t = 0;
  void* dblob = nullptr;
  unsigned long long* d_out = nullptr;
  unsigned long long* d_count = nullptr;
  uint32_t* d_mask_tc = nullptr;
  uint32_t* d_slow_tc = nullptr;
  DevPlan dp{};
  KernelFn kern = nullptr;
  // timing
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  uint64_t launches = 0, positions = 0;
};

namespace {

int upload_tconst(es_engine* e, const std::vector<uint32_t>& tc) {
  std::vector<uint32_t> mtc(std::max<uint32_t>(e->F, 1), 0), stc(std::max<uint32_t>(e->n_slow, 1), 0);
  for (uint32_t f = 0; f < e->F; ++f) mtc[f] = tc[e->mask_src[f]] % e->mask_mod[f];
  for (uint32_t j = 0; j < e->n_slow; ++j) stc[j] = tc[e->slow_order[j]];
  ES_HIP(hipMemcpyAsync(e->d_mask_tc, mtc.data(), mtc.size() * 4, hipMemcpyHostToDevice, e->stream));
  ES_HIP(hipMemcpyAsync(e->d_slow_tc, stc.data(), stc.size() * 4, hipMemcpyHostToDevice, e->stream));
  // the host vectors die here: make the copies complete before returning
  ES_HIP(hipStreamSynchronize(e->stream));
  return ES_OK;
}

int launch_range(es_engine* e, uint64_t o_lo, uint64_t o_hi, uint64_t idx_lo, uint64_t idx_hi) {
  if (o_hi <= o_lo) return ES_OK;
  const uint64_t u_lo = o_lo / e->S, u_hi = (o_hi + e->S - 1) / e->S;
  const uint64_t nstrip = u_hi - u_lo;
  uint64_t grid = (nstrip + e->threads - 1) / e->threads;
  if (grid > e->max_grid) grid = e->max_grid;
  if (grid == 0) grid = 1;
  hipEvent_t a = nullptr, b = nullptr;
  if (e->ev_used + 2 <= e->ev_pool.size()) {
    a = e->ev_pool[e->ev_used];
    b = e->ev_pool[e->ev_used + 1];
    e->ev_used += 2;
    ES_HIP(hipEventRecord(a, e->stream));
  }
  uint64_t cap = e->out_cap;
  void* args[] = {&e->dp, (void*)&u_lo, (void*)&u_hi, &o_lo, &o_hi, &idx_lo, &idx_hi,
                  &e->d_out, &e->d_count, &cap};
  ES_HIP(hipLaunchKernel((const void*)e->kern, dim3((uint32_t)grid), dim3(e->threads), args,
                         e->lds_bytes, e->stream));
  ES_HIP(hipGetLastError());
  if (b) {
    ES_HIP(hipEventRecord(b, e->stream));
    const uint64_t lo = std::max(idx_lo, o_lo * e->r0), hi = std::min(idx_hi, o_hi * e->r0);
    e->positions += hi > lo ? hi - lo : 0;
  }
  e->launches += 1;
  return ES_OK;
}

}  // namespace

extern "C" {

const char* es_last_error(void) { return g_err.c_str(); }

int es_device_query(int* n_devices, uint64_t* free_mib, uint64_t* total_mib, char* name,
                    uint32_t name_len) {
  int n = 0;
  ES_HIP(hipGetDeviceCount(&n));
  if (n_devices) *n_devices = n;
  if (n == 0) return fail(ES_ERR_DEVICE, "no HIP device visible");
  size_t fr = 0, tot = 0;
  ES_HIP(hipMemGetInfo(&fr, &tot));
  if (free_mib) *free_mib = fr >> 20;
  if (total_mib) *total_mib = tot >> 20;
  if (name && name_len) {
    int dev = 0;
    ES_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    ES_HIP(hipGetDeviceProperties(&prop, dev));
    std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  return ES_OK;
}

int es_engine_create(const es_plan_desc* d, es_engine** out) {
  if (!d || !out) return fail(ES_ERR_INVALID, "null argument");
  *out = nullptr;
  if (d->abi_version != ES_ABI_VERSION) return fail(ES_ERR_INVALID, "ABI version mismatch");
  if (d->n_rings < 1 || !d->ring_mod || !d->ring_radix || !d->term_lo || !d->term_hi)
    return fail(ES_ERR_INVALID, "plan needs at least one ring (pass a unit ring for an empty wheel)");
  const u128 M = mk128(d->M_lo, d->M_hi);
  if (M == 0 || (M >> 126) != 0) return fail(ES_ERR_INVALID, "wheel modulus must be in [1, 2^126)");
  const uint32_t w = d->n_rings;
  u128 prodq = 1, R = 1;
  std::vector<uint32_t> term_off(w + 1, 0);
  for (uint32_t j = 0; j < w; ++j) {
    if (d->ring_mod[j] == 0 || d->ring_radix[j] == 0 || d->ring_radix[j] > d->ring_mod[j])
      return fail(ES_ERR_INVALID, "bad ring modulus/radix");
    prodq *= d->ring_mod[j];
    R *= d->ring_radix[j];
    if ((R >> 62) != 0) return fail(ES_ERR_INVALID, "too many residues per block (R >= 2^62)");
    term_off[j + 1] = term_off[j] + d->ring_radix[j];
  }
  if (prodq != M) return fail(ES_ERR_INVALID, "ring moduli do not multiply to M");
  for (uint32_t f = 0; f < d->n_filters; ++f)
    if (d->filt_q[f] < 2 || d->filt_q[f] >= (1u << 24))
      return fail(ES_ERR_LIMIT, "filter modulus outside [2, 2^24)");
  if (d->n_masks > (uint32_t)kMaxMasks) return fail(ES_ERR_LIMIT, "too many word masks (max 12)");

  es_engine* e = new es_engine();
  std::unique_ptr<es_engine> guard(e);
  e->k = d->k;
  e->w = w;
  e->M = M;
  e->R = R;
  e->q0 = d->ring_mod[0];
  e->r0 = d->ring_radix[0];
  e->M0 = M / e->q0;
  e->W = (e->r0 + 31) / 32;
  e->threads = d->threads ? d->threads : 256;
  if (e->threads != 64 && e->threads != 128 && e->threads != 256 && e->threads != 512 &&
      e->threads != 1024)
    return fail(ES_ERR_INVALID, "threads must be 64, 128, 256, 512 or 1024");
  e->out_cap = d->out_cap ? d->out_cap : (1ull << 20);
  e->radix.assign(d->ring_radix, d->ring_radix + w);
  e->term_off = term_off;
  e->terms.resize(term_off[w]);
  for (uint32_t i = 0; i < term_off[w]; ++i) {
    e->terms[i] = mk128(d->term_lo[i], d->term_hi[i]);
    if (e->terms[i] >= M) return fail(ES_ERR_INVALID, "CRT term not reduced mod M");
  }
  e->n_filters = d->n_filters;
  e->filt_q.assign(d->filt_q, d->filt_q + d->n_filters);
  e->n_slow = d->n_slow;
  e->slow_order.assign(d->slow_order, d->slow_order + d->n_slow);
  for (uint32_t j : e->slow_order)
    if (j >= d->n_filters) return fail(ES_ERR_INVALID, "slow_order index out of range");

  // ---- ring 0: e values, sorted slot order
  const uint32_t r0 = e->r0, q0 = e->q0;
  std::vector<uint32_t> e_of(r0), order(r0);
  for (uint32_t i = 0; i < r0; ++i) {
    const u128 tv = e->terms[i];
    if (tv % e->M0 != 0) return fail(ES_ERR_INVALID, "ring-0 term is not a multiple of M/q0");
    e_of[i] = (uint32_t)(tv / e->M0);
    order[i] = i;
  }
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return e_of[a] < e_of[b]; });
  std::vector<uint32_t> e_sorted(r0), inner_ids(r0);
  for (uint32_t i = 0; i < r0; ++i) {
    e_sorted[i] = e_of[order[i]];
    inner_ids[i] = order[i];
  }
  for (uint32_t i = 1; i < r0; ++i)
    if (e_sorted[i] == e_sorted[i - 1]) return fail(ES_ERR_INVALID, "repeated ring-0 residue");

  // ---- strip rings 1..m and base rings m+1..w-1
  const uint32_t target = d->strip_target ? d->strip_target : 512;
  uint32_t m = 0;
  uint64_t S = 1;
  while (1 + m < w) {
    const uint64_t nxt = S * d->ring_radix[1 + m];
    if (m > 0 && nxt > target) break;
    if (nxt > (1u << 24)) break;
    S = nxt;
    ++m;
  }
  e->S = (uint32_t)S;
  e->R_outer = (uint64_t)(R / r0);
  e->strips = e->R_outer / e->S;
  const uint32_t nb = w - 1 - m;

  // ---- masks: (mask index f) -> rows of 2W words in LDS
  const uint32_t W = e->W;
  const uint32_t budget_words = (d->lds_budget ? d->lds_budget : 65536u) / 4;
  std::vector<uint32_t> lds;  // image
  std::vector<uint32_t> mrow, mmod, msrc;
  for (uint32_t f = 0; f < d->n_masks; ++f) {
    const uint32_t mq = d->mask_mod[f], src = d->mask_src[f];
    if (src >= d->n_filters || mq < 2 || d->filt_q[src] % mq != 0)
      return fail(ES_ERR_INVALID, "mask modulus must divide its filter's modulus");
    const size_t need = lds.size() + (size_t)2 * W * mq + q0 + 2;
    if (need > budget_words) break;  // later masks do not fit: they stay exact-only
    size_t off = (lds.size() + 1) & ~size_t(1);
    lds.resize(off + (size_t)2 * W * mq, 0);
    const uint32_t* bm = d->mask_bits + d->mask_off[f];
    const uint32_t Mm = mod_u128(M, mq);
    const uint32_t corr = (mq - Mm) % mq;
    std::vector<uint32_t> t0m(r0);
    for (uint32_t i = 0; i < r0; ++i) t0m[i] = mod_u128(e->terms[order[i]], mq);
    for (uint32_t x = 0; x < mq; ++x) {
      for (uint32_t i = 0; i < r0; ++i) {
        const uint32_t a = (uint32_t)(((uint64_t)x + t0m[i]) % mq);
        const uint32_t b = (uint32_t)(((uint64_t)a + corr) % mq);
        const uint32_t wi = i >> 5, bit = 1u << (i & 31);
        uint32_t* row = &lds[off + 2 * ((size_t)x * W + wi)];
        if ((bm[a >> 5] >> (a & 31)) & 1u) row[0] |= bit;
        if ((bm[b >> 5] >> (b & 31)) & 1u) row[1] |= bit;
      }
    }
    mrow.push_back((uint32_t)off);
    mmod.push_back(mq);
    msrc.push_back(src);
  }
  e->F = (uint32_t)mmod.size();
  e->mask_mod = mmod;
  e->mask_src = msrc;
  // split table: split[h] = #{ sorted slots i : e_i < q0 - h }
  const uint32_t split_off = (uint32_t)lds.size();
  lds.resize(split_off + q0, 0);
  for (uint32_t h = 0; h < q0; ++h) {
    const uint32_t thr = q0 - h;
    lds[split_off + h] =
        (uint32_t)(std::lower_bound(e_sorted.begin(), e_sorted.end(), thr) - e_sorted.begin());
  }
  if (lds.size() > budget_words && lds.size() > 16384)
    return fail(ES_ERR_LIMIT, "ring-0 modulus too large for the LDS split table");
  e->lds_bytes = (uint32_t)(lds.size() * 4);

  // ---- base-ring terms in (h, l) form
  std::vector<uint32_t> b_radix(std::max<uint32_t>(nb, 1), 0), b_off(std::max<uint32_t>(nb, 1), 0), t_h;
  std::vector<uint64_t> t_l;
  for (uint32_t j = 0; j < nb; ++j) {
    const uint32_t ring = 1 + m + j;
    b_radix[j] = d->ring_radix[ring];
    b_off[j] = (uint32_t)t_h.size();
    for (uint32_t i = term_off[ring]; i < term_off[ring + 1]; ++i) {
      const u128 tv = e->terms[i];
      t_h.push_back((uint32_t)(tv / e->M0));
      const u128 lv = tv % e->M0;
      t_l.push_back((uint64_t)lv);
      t_l.push_back((uint64_t)(lv >> 64));
    }
  }
  if (t_h.empty()) { t_h.push_back(0); t_l.push_back(0); t_l.push_back(0); }

  // ---- strip table over rings 1..m (ring 1 fastest)
  const uint32_t F = e->F;
  std::vector<uint32_t> s_h(S), s_q(std::max<uint64_t>(S * F * 2, 1), 0);
  std::vector<uint64_t> s_l(2 * S);
  std::vector<uint32_t> dig(m + 1, 0);
  for (uint64_t sg = 0; sg < S; ++sg) {
    u128 B = 0;
    for (uint32_t j = 0; j < m; ++j) {
      B += e->terms[term_off[1 + j] + dig[j]];
      if (B >= M) B -= M;
    }
    s_h[sg] = (uint32_t)(B / e->M0);
    const u128 L = B % e->M0;
    s_l[2 * sg] = (uint64_t)L;
    s_l[2 * sg + 1] = (uint64_t)(L >> 64);
    for (uint32_t f = 0; f < F; ++f) {
      const uint32_t mq = mmod[f];
      const uint32_t qv = mod_u128(B, mq);
      const uint32_t Mm = mod_u128(M, mq);
      s_q[(sg * F + f) * 2] = qv;
      s_q[(sg * F + f) * 2 + 1] = (uint32_t)(((uint64_t)qv + mq - Mm) % mq);
    }
    for (uint32_t j = 0; j < m; ++j) {  // odometer
      if (++dig[j] < d->ring_radix[1 + j]) break;
      dig[j] = 0;
    }
  }

  // ---- mask pow tables
  std::vector<uint32_t> m_pow(std::max<uint32_t>(4 * F, 4), 0), m_row(std::max<uint32_t>(F, 1), 0),
      m_mod(std::max<uint32_t>(F, 1), 1);
  for (uint32_t f = 0; f < F; ++f) {
    const uint32_t mq = mmod[f];
    m_pow[4 * f] = mod_u128(e->M0, mq);
    m_pow[4 * f + 1] = mod_u128((u128)1 << 32, mq);
    m_pow[4 * f + 2] = mod_u128((u128)1 << 64, mq);
    m_pow[4 * f + 3] = mod_u128((u128)1 << 96, mq);
    m_row[f] = mrow[f];
    m_mod[f] = mq;
  }
  // ---- slow filters
  const uint32_t ns = e->n_slow;
  std::vector<uint32_t> sq(std::max<uint32_t>(ns, 1), 2), soff(std::max<uint32_t>(ns, 1), 0),
      spow(std::max<uint32_t>(3 * ns, 3), 0);
  for (uint32_t j = 0; j < ns; ++j) {
    const uint32_t fi = e->slow_order[j], q = d->filt_q[fi];
    sq[j] = q;
    soff[j] = d->filt_off[fi];
    if ((uint64_t)d->filt_off[fi] + (q + 31) / 32 > d->n_filt_words)
      return fail(ES_ERR_INVALID, "filter bitmap outside filt_bits");
    spow[3 * j] = mod_u128((u128)1 << 32, q);
    spow[3 * j + 1] = mod_u128((u128)1 << 64, q);
    spow[3 * j + 2] = mod_u128((u128)1 << 96, q);
  }
  std::vector<uint32_t> fbits(d->filt_bits, d->filt_bits + std::max<uint32_t>(d->n_filt_words, 0));
  if (fbits.empty()) fbits.push_back(0);

  // ---- device upload: one blob
  Blob bl;
  const size_t o_lds = bl.put(lds.data(), lds.size() * 4);
  const size_t o_brad = bl.put(b_radix.data(), b_radix.size() * 4);
  const size_t o_boff = bl.put(b_off.data(), b_off.size() * 4);
  const size_t o_th = bl.put(t_h.data(), t_h.size() * 4);
  const size_t o_tl = bl.put(t_l.data(), t_l.size() * 8);
  const size_t o_sh = bl.put(s_h.data(), s_h.size() * 4);
  const size_t o_sl = bl.put(s_l.data(), s_l.size() * 8);
  const size_t o_sq = bl.put(s_q.data(), s_q.size() * 4);
  const size_t o_mmod = bl.put(m_mod.data(), m_mod.size() * 4);
  const size_t o_mrow = bl.put(m_row.data(), m_row.size() * 4);
  const size_t o_mpow = bl.put(m_pow.data(), m_pow.size() * 4);
  const size_t o_mtc = bl.put(nullptr, std::max<uint32_t>(F, 1) * 4);
  const size_t o_e0 = bl.put(e_sorted.data(), e_sorted.size() * 4);
  const size_t o_ids = bl.put(inner_ids.data(), inner_ids.size() * 4);
  const size_t o_sq2 = bl.put(sq.data(), sq.size() * 4);
  const size_t o_soff = bl.put(soff.data(), soff.size() * 4);
  const size_t o_sbits = bl.put(fbits.data(), fbits.size() * 4);
  const size_t o_spow = bl.put(spow.data(), spow.size() * 4);
  const size_t o_stc = bl.put(nullptr, std::max<uint32_t>(ns, 1) * 4);
  const size_t o_cnt = bl.put(nullptr, 8);

  ES_HIP(hipGetDevice(&e->device));
  hipDeviceProp_t prop;
  ES_HIP(hipGetDeviceProperties(&prop, e->device));
  if (e->lds_bytes > prop.sharedMemPerBlock)
    return fail(ES_ERR_LIMIT, "LDS tables exceed the device's per-workgroup LDS");
  e->max_grid = (uint32_t)std::max(1, prop.multiProcessorCount) * (2048u / e->threads) * 2u;
  ES_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  ES_HIP(hipMalloc(&e->dblob, bl.bytes.size()));
  ES_HIP(hipMemcpy(e->dblob, bl.bytes.data(), bl.bytes.size(), hipMemcpyHostToDevice));
  ES_HIP(hipMalloc(&e->d_out, std::max<uint64_t>(e->out_cap, 1) * 8));
  uint8_t* base = (uint8_t*)e->dblob;
  e->d_count = (unsigned long long*)(base + o_cnt);
  e->d_mask_tc = (uint32_t*)(base + o_mtc);
  e->d_slow_tc = (uint32_t*)(base + o_stc);

  DevPlan& P = e->dp;
  P.M0_lo = (uint64_t)e->M0;
  P.M0_hi = (uint64_t)(e->M0 >> 64);
  P.q0 = q0;
  P.r0 = r0;
  P.W = W;
  P.S = e->S;
  P.nb = nb;
  P.n_slow = ns;
  P.lds_words = (uint32_t)lds.size();
  P.split_off = split_off;
  P.lds_src = (const uint32_t*)(base + o_lds);
  P.base_radix = (const uint32_t*)(base + o_brad);
  P.base_off = (const uint32_t*)(base + o_boff);
  P.term_h = (const uint32_t*)(base + o_th);
  P.term_l = (const uint64_t*)(base + o_tl);
  P.strip_h = (const uint32_t*)(base + o_sh);
  P.strip_l = (const uint64_t*)(base + o_sl);
  P.strip_q = (const uint32_t*)(base + o_sq);
  P.mask_mod = (const uint32_t*)(base + o_mmod);
  P.mask_row = (const uint32_t*)(base + o_mrow);
  P.mask_pow = (const uint32_t*)(base + o_mpow);
  P.mask_tc = e->d_mask_tc;
  P.e0 = (const uint32_t*)(base + o_e0);
  P.inner_ids = (const uint32_t*)(base + o_ids);
  P.slow_q = (const uint32_t*)(base + o_sq2);
  P.slow_off = (const uint32_t*)(base + o_soff);
  P.slow_bits = (const uint32_t*)(base + o_sbits);
  P.slow_pow = (const uint32_t*)(base + o_spow);
  P.slow_tc = e->d_slow_tc;
  e->kern = pick_kernel((int)F);
  if (!e->kern) return fail(ES_ERR_LIMIT, "no kernel for this mask count");
  if (e->lds_bytes > 65536)
    ES_HIP(hipFuncSetAttribute((const void*)e->kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)e->lds_bytes));
  e->launch_outer = std::max<uint64_t>(1, ((uint64_t)1 << 36) / r0);
  e->ev_pool.resize(512);
  for (auto& ev : e->ev_pool) ES_HIP(hipEventCreate(&ev));
  ES_HIP(hipMemsetAsync(e->d_count, 0, 8, e->stream));
  std::vector<uint32_t> tc0(std::max<uint32_t>(d->n_filters, 1), 0);
  int rc = upload_tconst(e, tc0);
  if (rc) return rc;
  ES_HIP(hipStreamSynchronize(e->stream));
  *out = guard.release();
  return ES_OK;
}

int es_engine_destroy(es_engine* e) {
  if (!e) return ES_OK;
  if (e->stream) hipStreamSynchronize(e->stream);
  for (auto& ev : e->ev_pool) hipEventDestroy(ev);
  if (e->d_out) hipFree(e->d_out);
  if (e->dblob) hipFree(e->dblob);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return ES_OK;
}

int es_engine_info_get(const es_engine* e, es_engine_info* info) {
  if (!e || !info) return fail(ES_ERR_INVALID, "null argument");
  info->R_lo = (uint64_t)e->R;
  info->R_hi = (uint64_t)(e->R >> 64);
  info->R_outer = e->R_outer;
  info->strips = e->strips;
  info->r0 = e->r0;
  info->q0 = e->q0;
  info->strip_len = e->S;
  info->n_masks_used = e->F;
  info->lds_bytes = e->lds_bytes;
  info->words = e->W;
  info->threads = e->threads;
  info->device = (uint32_t)e->device;
  return ES_OK;
}

int es_engine_set_tconst(es_engine* e, const uint32_t* tconst, uint32_t n) {
  if (!e || (!tconst && n)) return fail(ES_ERR_INVALID, "null argument");
  if (n != e->n_filters) return fail(ES_ERR_INVALID, "tconst length must equal the filter count");
  std::vector<uint32_t> tc(tconst, tconst + n);
  tc.resize(std::max<uint32_t>(n, 1), 0);
  for (uint32_t j = 0; j < n; ++j)
    if (tc[j] >= e->filt_q[j]) return fail(ES_ERR_INVALID, "tconst entry not reduced mod q");
  return upload_tconst(e, tc);
}

int es_engine_set_block(es_engine* e, uint64_t t) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  std::vector<uint32_t> tc(std::max<uint32_t>(e->n_filters, 1), 0);
  for (uint32_t j = 0; j < e->n_filters; ++j) {
    const uint32_t q = e->filt_q[j];
    tc[j] = (uint32_t)(((uint64_t)(t % q) * mod_u128(e->M, q)) % q);
  }
  e->t = t;
  return upload_tconst(e, tc);
}

int es_engine_reset(es_engine* e) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  ES_HIP(hipMemsetAsync(e->d_count, 0, 8, e->stream));
  return ES_OK;
}

int es_engine_scan_window(es_engine* e, uint64_t idx_lo, uint64_t idx_hi) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  if (idx_hi > (uint64_t)e->R || idx_lo > idx_hi)
    return fail(ES_ERR_INVALID, "window outside [0, R)");
  if (idx_lo == idx_hi) return ES_OK;
  const uint64_t o_lo = idx_lo / e->r0, o_hi = (idx_hi + e->r0 - 1) / e->r0;
  for (uint64_t a = o_lo; a < o_hi; a += e->launch_outer) {
    const uint64_t b = std::min(o_hi, a + e->launch_outer);
    int rc = launch_range(e, a, b, idx_lo, idx_hi);
    if (rc) return rc;
  }
  return ES_OK;
}

int es_engine_scan(es_engine* e, uint64_t ostart, uint64_t ocount) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  if (ostart > e->R_outer || ocount > e->R_outer - ostart)
    return fail(ES_ERR_INVALID, "outer range outside [0, R_outer)");
  return es_engine_scan_window(e, ostart * e->r0, (ostart + ocount) * e->r0);
}

int es_engine_sync(es_engine* e) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  ES_HIP(hipStreamSynchronize(e->stream));
  return ES_OK;
}

int es_engine_count(es_engine* e, uint64_t* n) {
  if (!e || !n) return fail(ES_ERR_INVALID, "null argument");
  unsigned long long c = 0;
  ES_HIP(hipMemcpyAsync(&c, e->d_count, 8, hipMemcpyDeviceToHost, e->stream));
  ES_HIP(hipStreamSynchronize(e->stream));
  *n = c;
  return ES_OK;
}

int es_engine_read(es_engine* e, uint64_t offset, uint64_t count, uint64_t* ids) {
  if (!e || (!ids && count)) return fail(ES_ERR_INVALID, "null argument");
  if (offset > e->out_cap || count > e->out_cap - offset)
    return fail(ES_ERR_INVALID, "read outside the survivor buffer");
  if (count)
    ES_HIP(hipMemcpyAsync(ids, e->d_out + offset, count * 8, hipMemcpyDeviceToHost, e->stream));
  ES_HIP(hipStreamSynchronize(e->stream));
  return ES_OK;
}

int es_engine_collect(es_engine* e, uint64_t* ids, uint64_t cap, uint64_t* n, uint64_t min_s[2],
                      int* has_min) {
  if (!e || !n) return fail(ES_ERR_INVALID, "null argument");
  uint64_t c = 0;
  int rc = es_engine_count(e, &c);
  if (rc) return rc;
  *n = c;
  if (has_min) *has_min = 0;
  if (c > e->out_cap) return fail(ES_ERR_OVERFLOW, "survivor buffer overflow");
  std::vector<uint64_t> buf(c);
  if (c) {
    rc = es_engine_read(e, 0, c, buf.data());
    if (rc) return rc;
  }
  if (ids) std::memcpy(ids, buf.data(), std::min<uint64_t>(c, cap) * 8);
  // least s with t*M + s > k + 1
  bool have = false;
  u128 best = 0;
  for (uint64_t v : buf) {
    uint64_t idx = v;
    u128 s = 0;
    for (uint32_t j = 0; j < e->w; ++j) {
      const uint64_t dj = idx % e->radix[j];
      idx /= e->radix[j];
      s += e->terms[e->term_off[j] + dj];
      if (s >= e->M) s -= e->M;
    }
    if (e->t == 0 && s <= (u128)e->k + 1) continue;
    if (!have || s < best) {
      best = s;
      have = true;
    }
  }
  if (have && min_s) {
    min_s[0] = (uint64_t)best;
    min_s[1] = (uint64_t)(best >> 64);
  }
  if (has_min) *has_min = have ? 1 : 0;
  return ES_OK;
}

int es_engine_kernel_time(es_engine* e, double* ms, uint64_t* launches, uint64_t* positions) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  ES_HIP(hipStreamSynchronize(e->stream));
  double tot = 0.0;
  for (size_t i = 0; i + 1 < e->ev_used; i += 2) {
    float t = 0.f;
    ES_HIP(hipEventElapsedTime(&t, e->ev_pool[i], e->ev_pool[i + 1]));
    tot += t;
  }
  const uint64_t timed = e->ev_used / 2;
  if (ms) *ms = tot;
  if (launches) *launches = timed;
  if (positions) *positions = e->positions;
  e->ev_used = 0;
  e->launches = 0;
  e->positions = 0;
  return ES_OK;
}

}  // extern "C"
