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
// iff h + e_i >= q0, so the ring-0 wrap flags of all r0 positions of an outer index are one
// table row indexed by the small integer h, never a 128-bit compare.
//
// Strips.  Rings 1..m form a strip of S = r_1*...*r_m outer indices that every lane of a
// wave walks in the same order, so the strip's partial sums (H, L) and their residues mod
// each mask modulus are wave-uniform rows (scalar loads).  Rings m+1.. are the lane's strip
// base u, decoded once per strip.  Per outer index a lane does one 128-bit compare (carry of
// l into h), a few 32-bit ops, and for each word-mask prefilter one aligned 8-byte LDS load:
// table f keeps the r0-bit masks of phase x = (b + tM) mod m_f (bit i: n passes the prefilter
// at ring-0 slot i) at row x * (-M)^-1 mod m_f, as two words: the mask of x and that of
// x - M (the phase of a slot whose inner term wraps past M), which is the next row's first
// word.  Rows are stored twice over (rows [0, 2m)) so P(BQ_u) + P(Q_sigma) + wrap needs no
// reduction, and each lane keeps the LDS byte address of its strip-base row, so a mask costs
// one add, one ds_read and one 3-input AND; the wrap row selects between the two ANDed words
// once.
//
// Exactness.  Prefilters are necessary conditions (low base-p digits of a filter prime), so
// they only discard positions the full predicate rejects.  Every surviving position is then
// re-tested with the full admissibility tables of every remaining filter from its exact
// 128-bit residue s -- the reference predicate (sieve/gpu_sieve_v9.py:304-322).  Mask words
// that still have candidates are queued per wave and re-tested 64 at a time.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "es_common.h"
#include "es_sieve_dev.h"

namespace es {

std::string& last_error() {
  thread_local std::string msg;
  return msg;
}

int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

}  // namespace es

using es::fail;
using es::mk128;
using es::mod_u128;
using es::inv_mod;

using namespace es_dev;

namespace {

using namespace es_dev;

KernelFn pick_kernel(int F, uint32_t W, uint32_t G, uint32_t C) {
  if (F < 0 || F > kMaxMasks) return nullptr;
  if (F < kPartFirstF[1]) return pick_kernel_p0(F, W, G, C);
  if (F < kPartFirstF[2]) return pick_kernel_p1(F, W, G, C);
  if (F < kPartFirstF[3]) return pick_kernel_p2(F, W, G, C);
  return pick_kernel_p3(F, W, G, C);
}
RowsFn pick_rows(int F, uint32_t C, uint32_t W, bool wave8) {
  if (F < 1 || F > kRowsMaxMasks) return nullptr;
  if (F < kPartFirstF[1]) return pick_rows_p0(F, C, W, wave8);
  if (F < kPartFirstF[2]) return pick_rows_p1(F, C, W, wave8);
  if (F < kPartFirstF[3]) return pick_rows_p2(F, C, W, wave8);
  return pick_rows_p3(F, C, W, wave8);
}
RowsBatchFn pick_rows_batch(int F, uint32_t C, uint32_t W, bool wave8) {
  if (F < 1 || F > kRowsMaxMasks) return nullptr;
  if (F < kPartFirstF[1]) return pick_rows_batch_p0(F, C, W, wave8);
  if (F < kPartFirstF[2]) return pick_rows_batch_p1(F, C, W, wave8);
  if (F < kPartFirstF[3]) return pick_rows_batch_p2(F, C, W, wave8);
  return pick_rows_batch_p3(F, C, W, wave8);
}

// Lane slots of the strip rows for the row-lane layout: G groups of 64 slots, i.e. 2G
// half-waves of 32.  bank[sg * F + f] is the bank pair (row mod 32) of strip row sg in mask
// table f; a half-wave costs, per table, its most crowded bank pair (the cycles one
// ds_read_b64 of those 32 lanes takes).  Rows are placed greedily (each to the open half-wave
// it crowds least, in a seeded shuffled order), then pairs of slots in different half-waves
// are swapped while the swap does not raise the two half-waves' cost.  The first slot of a
// pair is drawn from a half-wave above the floor of two lanes on the most crowded pair of
// every table (~31 live rows over 32 pairs rarely allow fewer), on a row that sits on such a
// pair; the search stops once no half-wave is above it.  (Random pairs, the earlier search,
// left wheels with long strips at 4.4-4.5 cycles per read where this reaches 4.0 in the same
// number of swaps.)  Empty slots repeat a row of their half-wave (the same address: no extra
// bank load) and carry bit 31 (no output).
std::vector<uint32_t> group_strip_rows(const std::vector<uint32_t>& bank, uint32_t S, uint32_t F,
                                       uint32_t G) {
  const uint32_t H = 2 * G;                      // half-waves
  const uint32_t FF = std::max<uint32_t>(F, 1);
  // per (half-wave, table): bank-pair loads, the largest load and how many pairs carry it;
  // per half-wave: its cost, the sum of the largest loads over the tables
  std::vector<uint8_t> cnt((size_t)H * FF * 32, 0), mx((size_t)H * FF, 0), nmx((size_t)H * FF, 32);
  std::vector<uint32_t> hcost(H, 0);
  std::vector<uint32_t> slot((size_t)H * 32, UINT32_MAX), fill(H, 0);
  auto bin = [&](uint32_t h, uint32_t f, uint32_t b) -> uint8_t& {
    return cnt[((size_t)h * FF + f) * 32 + b];
  };
  auto inc = [&](uint32_t h, uint32_t f, uint32_t b) {
    const size_t k = (size_t)h * FF + f;
    const uint8_t c = ++bin(h, f, b);
    if (c > mx[k]) {
      mx[k] = c;
      nmx[k] = 1;
      ++hcost[h];
    } else if (c == mx[k]) {
      ++nmx[k];
    }
  };
  auto dec = [&](uint32_t h, uint32_t f, uint32_t b) {
    const size_t k = (size_t)h * FF + f;
    const uint8_t c = bin(h, f, b)--;
    if (c == mx[k] && --nmx[k] == 0) {   // the only most crowded pair lost a lane: recount
      mx[k] = c - 1;
      --hcost[h];
      uint8_t n = 0;
      for (uint32_t x = 0; x < 32; ++x) n += bin(h, f, x) == mx[k];
      nmx[k] = n;
    }
  };
  uint64_t rng = 0x9e3779b97f4a7c15ull;
  auto next = [&]() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
  };
  std::vector<uint32_t> order(S);
  for (uint32_t i = 0; i < S; ++i) order[i] = i;
  for (uint32_t i = S; i > 1; --i) std::swap(order[i - 1], order[next() % i]);
  for (uint32_t r : order) {
    uint32_t best_h = UINT32_MAX, best_add = UINT32_MAX;
    for (uint32_t h = 0; h < H && best_add; ++h) {
      if (fill[h] >= 32) continue;
      uint32_t add = 0;
      for (uint32_t f = 0; f < F; ++f)
        add += bin(h, f, bank[(size_t)r * F + f]) + 1u > mx[(size_t)h * FF + f] ? 1u : 0u;
      if (add < best_add) {
        best_add = add;
        best_h = h;
      }
    }
    slot[(size_t)best_h * 32 + fill[best_h]++] = r;
    for (uint32_t f = 0; f < F; ++f) inc(best_h, f, bank[(size_t)r * F + f]);
  }
  auto move = [&](uint32_t r, uint32_t from, uint32_t to) {
    if (r == UINT32_MAX) return;
    for (uint32_t f = 0; f < F; ++f) {
      dec(from, f, bank[(size_t)r * F + f]);
      inc(to, f, bank[(size_t)r * F + f]);
    }
  };
  const uint32_t floor_cost = 2 * F;
  auto above_floor = [&]() {
    uint32_t n = 0;
    for (uint32_t h = 0; h < H; ++h) n += hcost[h] > floor_cost;
    return n;
  };
  // swaps: at most 120 per strip row (200,000 in all); stop once every half-wave is at the
  // floor, or after 20,000 swaps without a lower total cost
  const uint64_t iters = std::min<uint64_t>(200000, (uint64_t)S * 120);
  const uint32_t n_slots = H * 32;
  uint64_t total = 0, last_gain = 0;
  for (uint32_t h = 0; h < H; ++h) total += hcost[h];
  for (uint64_t it = 0; it < iters && H > 1; ++it) {
    if ((it & 1023u) == 0 && (above_floor() == 0 || it - last_gain > 20000)) break;
    // first slot: a half-wave above the floor (a few draws), a row of it on a pair loaded
    // above two in some table; else any slot
    uint32_t a = (uint32_t)(next() % n_slots);
    {
      uint32_t h = (uint32_t)(next() % H), tries = 0;
      while (hcost[h] <= floor_cost && ++tries < 8) h = (uint32_t)(next() % H);
      if (hcost[h] > floor_cost) {
        const uint32_t s0 = (uint32_t)(next() % 32);
        for (uint32_t k = 0; k < 32; ++k) {
          const uint32_t r = slot[(size_t)h * 32 + (s0 + k) % 32];
          if (r == UINT32_MAX) continue;
          bool hot = false;
          for (uint32_t f = 0; f < F && !hot; ++f) {
            const uint8_t c = bin(h, f, bank[(size_t)r * F + f]);
            hot = c > 2 && c == mx[(size_t)h * FF + f];
          }
          if (hot) {
            a = h * 32 + (s0 + k) % 32;
            break;
          }
        }
      }
    }
    const uint32_t b = (uint32_t)(next() % n_slots);
    const uint32_t ha = a / 32, hb = b / 32;
    const uint32_t ra = slot[a], rb = slot[b];
    if (ha == hb || (ra == UINT32_MAX && rb == UINT32_MAX)) continue;
    const uint32_t before = hcost[ha] + hcost[hb];
    move(ra, ha, hb);
    move(rb, hb, ha);
    const uint32_t after = hcost[ha] + hcost[hb];
    if (after <= before) {
      slot[a] = rb;
      slot[b] = ra;
      if (after < before) {
        total -= before - after;
        last_gain = it;
      }
    } else {
      move(ra, hb, ha);
      move(rb, ha, hb);
    }
  }
  std::vector<uint32_t> perm((size_t)G * 64, 0);
  for (uint32_t h = 0; h < H; ++h) {
    uint32_t any = UINT32_MAX;
    for (uint32_t i = 0; i < 32 && any == UINT32_MAX; ++i) any = slot[(size_t)h * 32 + i];
    if (any == UINT32_MAX) any = 0;                // an empty half-wave (S small): row 0
    for (uint32_t i = 0; i < 32; ++i) {
      const uint32_t r = slot[(size_t)h * 32 + i];
      perm[(size_t)h * 32 + i] = r == UINT32_MAX ? (any | 0x80000000u) : r;
    }
  }
  return perm;
}

GridBatchFn pick_grid_batch(int F, uint32_t W, uint32_t G, uint32_t C) {
  if (F < 0 || F > kMaxMasks) return nullptr;
  if (F < kPartFirstF[1]) return pick_grid_batch_p0(F, W, G, C);
  if (F < kPartFirstF[2]) return pick_grid_batch_p1(F, W, G, C);
  if (F < kPartFirstF[3]) return pick_grid_batch_p2(F, W, G, C);
  return pick_grid_batch_p3(F, W, G, C);
}
struct Blob {
  std::vector<uint8_t> bytes;
  size_t put(const void* p, size_t n) {
    const size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + std::max<size_t>(n, 16), 0);
    if (n && p) std::memcpy(bytes.data() + off, p, n);
    return off;
  }
};

}  // namespace

struct es_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t k = 0, w = 0, r0 = 1, q0 = 1, W = 1, S = 1, F = 0, G = 0, threads = 256;
  uint32_t n_filters = 0, n_slow = 0, lds_bytes = 0, max_grid = 1;
  u128 M = 1, M0 = 1, R = 1;
  uint64_t R_outer = 1, strips = 1, out_cap = 1u << 20, launch_outer = 1;
  uint32_t lead = 1;   // > 1: the reference's ring j leads (product of the radices before it)
  uint32_t lead_rings = 0;  // j: the kernel's rings 1..j are the reference's rings 0..j-1
  uint32_t unit = 1;   // positions per es_engine_scan outer unit: the reference's ring-0 radix
  uint32_t kind = 0;   // kernel instance C: 1 byte carry table, 2 lead block idx form
  std::vector<uint32_t> filt_q, slow_order, mask_mod, mask_src, radix;
  std::vector<uint32_t> mask_info, slow_info;  // host copies of the device records
  std::vector<u128> terms;                      // all ring terms, ring-major (host idx -> s)
  std::vector<uint32_t> term_off;
  uint64_t t = 0;
  bool t_known = true;       // false after es_engine_set_tconst: the minimum needs the block
  uint64_t last_count = 0;   // survivor count read since the last launch or reset
  bool count_valid = false;
  void* dblob = nullptr;
  unsigned long long* d_out = nullptr;    // [out_cap] IDs, then [out_cap][2] residues
  unsigned long long* d_count = nullptr;  // count, 1 + slot of the minimum, its floor
  uint32_t* d_mask_info = nullptr;
  uint32_t* d_slow_info = nullptr;
  HotPlan hp{};
  ColdPlan cold{};
  KernelFn kern = nullptr;
  bool rows = false;         // row-lane layout (lanes over strip rows, waves over steps)
  uint32_t V = 1, Gl = 0;    // row-lane: steps per unit, lane groups of 64 strip rows
  RowPlan rp{};
  RowPlan* d_rp = nullptr;   // device copy of rp (the row-lane kernel takes a pointer)
  RowsFn rows_kern = nullptr;
  bool wave8 = false;        // row-lane kernel instance sized for 8 waves per SIMD (else 6)
  // with wave8: the 6-wave instance (no spills, so no scratch memory) and its workgroup size
  // and LDS bytes, for the scan that allocates nothing
  RowsFn rows_kern6 = nullptr;
  uint32_t threads6 = 0, lds_bytes6 = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  uint64_t launches = 0, positions = 0;
  uint64_t floor = 0;        // host copy of d_count[2]
  // es_engine_scan_windows: two result slots (set up on first use), so that window i + 1 runs
  // while the host reads window i
  struct Slot {
    unsigned long long* d_out = nullptr;  // [out_cap] IDs, then [out_cap][2] residues
    unsigned long long* d_ctr = nullptr;  // count, 1 + slot of the minimum, its floor
    ColdPlan* d_cold = nullptr;           // the exact test's record with this slot's buffers
    RowPlan* d_rp = nullptr;              // the row-lane plan pointing at d_cold
    hipEvent_t done = nullptr;            // the slot's header is written
  };
  Slot slot[2];
  void* slot_mem = nullptr;
  uint64_t* h_hdr = nullptr;              // pinned, mapped [2][4] headers (count, least s lo,
                                          // hi, 1 if there is a minimum), then [4] counter set-up
  uint64_t* h_hdr_dev = nullptr;          // h_hdr as the device addresses it
  hipStream_t copy_stream = nullptr;      // survivor read-back beside the sieve stream
};

namespace {

// Block constants, and the least residue s that counts for the minimum: n = t*M + s > k + 1,
// i.e. s >= k + 2 in block 0 and every s in later blocks (0 as well while the block is unknown).
int upload_tconst(es_engine* e, const std::vector<uint32_t>& tc) {
  for (uint32_t f = 0; f < e->F; ++f)
    e->mask_info[kMaskInfo * f + 7] = tc[e->mask_src[f]] % e->mask_mod[f];
  for (uint32_t j = 0; j < e->n_slow; ++j) e->slow_info[kSlowInfo * j + 5] = tc[e->slow_order[j]];
  e->floor = (e->t_known && e->t == 0) ? (uint64_t)e->k + 2 : 0;
  ES_HIP(hipMemcpyAsync(e->d_mask_info, e->mask_info.data(), e->mask_info.size() * 4,
                        hipMemcpyHostToDevice, e->stream));
  ES_HIP(hipMemcpyAsync(e->d_slow_info, e->slow_info.data(), e->slow_info.size() * 4,
                        hipMemcpyHostToDevice, e->stream));
  ES_HIP(hipMemcpyAsync(e->d_count + 2, &e->floor, 8, hipMemcpyHostToDevice, e->stream));
  ES_HIP(hipStreamSynchronize(e->stream));  // host records may change right after
  return ES_OK;
}

// Kernel outer indices [o_lo, o_hi) whose positions include every idx in [idx_lo, idx_hi):
// idx / r0 in the reference order; with a lead block, the whole runs of lead * r0 idx
// (lead outer indices each) that meet the window.  The exact test drops positions outside it.
void outer_range(const es_engine* e, uint64_t idx_lo, uint64_t idx_hi, uint64_t* o_lo,
                 uint64_t* o_hi) {
  const uint64_t per = (uint64_t)e->lead * e->r0;
  *o_lo = idx_lo / per * e->lead;
  *o_hi = idx_hi > idx_lo ? ((idx_hi - 1) / per + 1) * e->lead : *o_lo;
}

// One kernel launch over outer indices [o_lo, o_hi) on stream st; a non-null `cold` replaces
// the strip-lane plan's exact-test record, and a non-null `rows_plan` is the device copy of
// the row-lane plan to use (the static-scratch scan).  Strip-lane: one lane per strip;
// row-lane: one wave per unit (lane group, superstrip).
int launch_kernel(es_engine* e, hipStream_t st, const ColdPlan* cold, const RowPlan* rows_plan,
                  uint64_t o_lo, uint64_t o_hi, uint64_t idx_lo, uint64_t idx_hi,
                  bool no_scratch = false) {
  if (e->rows) {
    // one wave per unit; superstrips [w_lo, w_hi) in launches of at most max_grid workgroups;
    // no_scratch: the 6-wave instance where the engine runs the 8-wave one
    const bool six = no_scratch && e->wave8;
    const void* fn = six ? (const void*)e->rows_kern6 : (const void*)e->rows_kern;
    const uint32_t threads = six ? e->threads6 : e->threads;
    const uint32_t lds = six ? e->lds_bytes6 : e->lds_bytes;
    const uint64_t span = (uint64_t)e->S * e->V;
    const uint64_t w_end = (o_hi + span - 1) / span;
    const uint32_t wpg = threads / 64;
    const uint64_t per = std::max<uint64_t>(1, (uint64_t)e->max_grid * wpg / e->Gl);
    const RowPlan* rp = rows_plan ? rows_plan : e->d_rp;
    for (uint64_t w_lo = o_lo / span; w_lo < w_end; w_lo += per) {
      uint64_t a = w_lo, b = std::min(w_end, w_lo + per);
      const uint64_t grid = ((uint64_t)e->Gl * (b - a) + wpg - 1) / wpg;
      void* args[] = {&rp, &a, &b, &o_lo, &o_hi, &idx_lo, &idx_hi};
      ES_HIP(hipLaunchKernel(fn, dim3((uint32_t)grid), dim3(threads), args, lds, st));
    }
  } else {
    uint64_t u_lo = o_lo / e->S, u_hi = (o_hi + e->S - 1) / e->S;
    const uint64_t grid = std::max<uint64_t>(
        1, std::min<uint64_t>((u_hi - u_lo + e->threads - 1) / e->threads, e->max_grid));
    HotPlan hp = e->hp;
    if (cold) hp.cold = cold;
    void* args[] = {&hp, &u_lo, &u_hi, &o_lo, &o_hi, &idx_lo, &idx_hi};
    ES_HIP(hipLaunchKernel((const void*)e->kern, dim3((uint32_t)grid), dim3(e->threads), args,
                           e->lds_bytes, st));
  }
  ES_HIP(hipGetLastError());
  return ES_OK;
}

// One timed launch on the engine's stream; `cold` / `rows_plan` as in launch_kernel (a result
// slot of es_engine_scan_windows), or null for the engine's own survivor buffer.
int launch_range(es_engine* e, uint64_t o_lo, uint64_t o_hi, uint64_t idx_lo, uint64_t idx_hi,
                 const ColdPlan* cold = nullptr, const RowPlan* rows_plan = nullptr) {
  if (!cold) e->count_valid = false;
  if (o_hi <= o_lo) return ES_OK;
  hipEvent_t a = nullptr, b = nullptr;
  if (e->ev_used + 2 <= e->ev_pool.size()) {
    a = e->ev_pool[e->ev_used];
    b = e->ev_pool[e->ev_used + 1];
    e->ev_used += 2;
    ES_HIP(hipEventRecord(a, e->stream));
  }
  const int rc = launch_kernel(e, e->stream, cold, rows_plan, o_lo, o_hi, idx_lo, idx_hi);
  if (rc) return rc;
  if (b) {
    ES_HIP(hipEventRecord(b, e->stream));
    // idx covered by outer indices [o_lo, o_hi) (multiples of lead when rings are swapped)
    const uint64_t per = (uint64_t)e->lead * e->r0;
    const uint64_t lo = std::max(idx_lo, o_lo / e->lead * per),
                   hi = std::min(idx_hi, o_hi / e->lead * per);
    e->positions += hi > lo ? hi - lo : 0;
  }
  e->launches += 1;
  return ES_OK;
}

// Positions [idx_lo, idx_hi) of the current block in equal launches of at most launch_outer
// outer indices (no short tail launch), each a multiple of lead so it covers whole runs of idx.
int scan_range(es_engine* e, uint64_t idx_lo, uint64_t idx_hi, const ColdPlan* cold,
               const RowPlan* rows_plan) {
  if (idx_lo == idx_hi) return ES_OK;
  uint64_t o_lo, o_hi;
  outer_range(e, idx_lo, idx_hi, &o_lo, &o_hi);
  const uint64_t n = (o_hi - o_lo + e->launch_outer - 1) / e->launch_outer;
  uint64_t per = (o_hi - o_lo + n - 1) / n;
  per = (per + e->lead - 1) / e->lead * e->lead;
  for (uint64_t a = o_lo; a < o_hi; a += per) {
    const uint64_t b = std::min(o_hi, a + per);
    const int rc = launch_range(e, a, b, idx_lo, idx_hi, cold, rows_plan);
    if (rc) return rc;
  }
  return ES_OK;
}

// A result slot's header after its window, written straight to pinned host memory:
// {count (may exceed the slot's capacity), least residue s lo, hi, 1 if there is a minimum};
// the counter and the minimum's slot then start again at zero for the slot's next window (its
// floor stays).
__global__ void es_slot_finish_kernel(unsigned long long* __restrict__ ctr,
                                      const unsigned long long* __restrict__ out_s,
                                      uint64_t* __restrict__ hdr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const uint64_t best = ctr[1];
    hdr[0] = ctr[0];
    hdr[1] = best ? out_s[2 * (best - 1)] : 0;
    hdr[2] = best ? out_s[2 * (best - 1) + 1] : 0;
    hdr[3] = best ? 1 : 0;
    __threadfence_system();
    ctr[0] = 0;
    ctr[1] = 0;
  }
}

// Both result slots of es_engine_scan_windows: device buffers like the engine's own (IDs and
// residue records for out_cap survivors, counters), the exact test's record and the row-lane
// plan pointing at them, the pinned host headers (mapped for the device) and the read-back
// stream.
int setup_slots(es_engine* e) {
  if (e->slot_mem) return ES_OK;
  const uint64_t cap = std::max<uint64_t>(e->out_cap, 1);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_ctr = al(cap * 24), o_cold = o_ctr + 256, o_rp = o_cold + al(sizeof(ColdPlan)),
               per = o_rp + al(sizeof(RowPlan));
  void* mem = nullptr;
  ES_HIP(hipMalloc(&mem, 2 * per));
  e->slot_mem = mem;
  ES_HIP(hipHostMalloc((void**)&e->h_hdr, 16 * sizeof(uint64_t),
                       hipHostMallocMapped | hipHostMallocCoherent));
  ES_HIP(hipHostGetDevicePointer((void**)&e->h_hdr_dev, e->h_hdr, 0));
  if (!e->h_hdr_dev) return fail(ES_ERR_DEVICE, "pinned header has no device address");
  ES_HIP(hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking));
  for (int i = 0; i < 2; ++i) {
    es_engine::Slot& sl = e->slot[i];
    uint8_t* b = (uint8_t*)mem + i * per;
    sl.d_out = (unsigned long long*)b;
    sl.d_ctr = (unsigned long long*)(b + o_ctr);
    sl.d_cold = (ColdPlan*)(b + o_cold);
    sl.d_rp = (RowPlan*)(b + o_rp);
    ES_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    ColdPlan c = e->cold;
    c.out = sl.d_out;
    c.out_n = sl.d_ctr;
    c.out_s = sl.d_out + cap;
    c.out_cap = e->out_cap;
    ES_HIP(hipMemcpy(sl.d_cold, &c, sizeof(c), hipMemcpyHostToDevice));
    RowPlan q = e->rp;
    q.cold = sl.d_cold;
    ES_HIP(hipMemcpy(sl.d_rp, &q, sizeof(q), hipMemcpyHostToDevice));
  }
  return ES_OK;
}

}  // namespace

extern "C" {

const char* es_last_error(void) { return es::last_error().c_str(); }

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
    std::snprintf(name, name_len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount);
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
  if (d->n_filters && (!d->filt_q || !d->filt_off || !d->filt_bits))
    return fail(ES_ERR_INVALID, "filter arrays missing");
  for (uint32_t f = 0; f < d->n_filters; ++f) {
    if (d->filt_q[f] < 2 || d->filt_q[f] >= (1u << 24))
      return fail(ES_ERR_LIMIT, "filter modulus outside [2, 2^24)");
    if ((uint64_t)d->filt_off[f] + (d->filt_q[f] + 31) / 32 > d->n_filt_words)
      return fail(ES_ERR_INVALID, "filter bitmap outside filt_bits");
  }
  for (uint32_t j = 0; j < d->n_slow; ++j)
    if (d->slow_order[j] >= d->n_filters) return fail(ES_ERR_INVALID, "slow_order index out of range");
  for (uint32_t f = 0; f < d->n_masks; ++f) {
    const uint32_t mq = d->mask_mod[f], src = d->mask_src[f];
    if (src >= d->n_filters || mq < 2 || d->filt_q[src] % mq != 0)
      return fail(ES_ERR_INVALID, "mask modulus must divide its filter's modulus");
    if ((uint64_t)d->mask_off[f] + (mq + 31) / 32 > d->n_mask_words)
      return fail(ES_ERR_INVALID, "mask bitmap outside mask_bits");
  }

  std::unique_ptr<es_engine> guard(new es_engine());
  es_engine* e = guard.get();
  e->k = d->k;
  e->w = w;
  e->M = M;
  e->R = R;
  e->q0 = d->inner_mod ? d->inner_mod : d->ring_mod[0];
  if (e->q0 % d->ring_mod[0] != 0 || M % e->q0 != 0 || e->q0 >= (1u << 31))
    return fail(ES_ERR_INVALID, "inner_mod must divide M, be a multiple of ring_mod[0] and < 2^31");
  e->r0 = d->ring_radix[0];
  e->lead = d->lead_radix > 1 ? d->lead_radix : 1;
  if (e->lead > 1) {
    // the lead block: rings 1..lead_rings, whose radices multiply to lead_radix
    uint64_t L = 1;
    while (L < e->lead && 1 + e->lead_rings < w) L *= d->ring_radix[1 + e->lead_rings++];
    if (L != e->lead || d->inner_mod != 0)
      return fail(ES_ERR_INVALID, "lead_radix must be the product of ring_radix[1..j] for some j "
                                  "(and ring 0 not a digit split)");
  }
  e->unit = e->lead > 1 ? d->ring_radix[1] : e->r0;  // the reference's ring-0 radix
  e->M0 = M / e->q0;
  e->W = (e->r0 + 31) / 32;
  e->threads = d->threads;  // 0: chosen by occupancy once the tables are sized
  if (e->threads != 0 && e->threads != 64 && e->threads != 128 && e->threads != 256 &&
      e->threads != 512 && e->threads != 1024)
    return fail(ES_ERR_INVALID, "threads must be 0, 64, 128, 256, 512 or 1024");
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
  e->slow_order.assign(d->slow_order, d->slow_order + d->n_slow);

  // ---- ring 0: e values, sorted slot order
  const uint32_t r0 = e->r0, q0 = e->q0, W = e->W;
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
    inner_ids[i] = order[i] * e->lead;  // idx offset of the slot (ring-0 digit x lead)
  }
  for (uint32_t i = 1; i < r0; ++i)
    if (e_sorted[i] == e_sorted[i - 1]) return fail(ES_ERR_INVALID, "repeated ring-0 residue");

  // ---- strip rings 1..m and base rings m+1..w-1
  const uint32_t target = d->strip_target ? d->strip_target : 1024;
  uint32_t m = 0;
  uint64_t S = 1;
  // strip rings 1..m: the product of leading radices nearest the target in log scale (the
  // ring that crosses the target is taken when S * nxt < target^2)
  while (1 + m < w) {
    const uint64_t nxt = S * d->ring_radix[1 + m];
    if (nxt > (1u << 22) || nxt * e->W >= (1u << kQueueIndexBits)) break;
    if (m > 0 && m >= e->lead_rings && nxt > target) {  // the lead block is always whole
      if (nxt * S < (uint64_t)target * target) {
        S = nxt;
        ++m;
      }
      break;
    }
    S = nxt;
    ++m;
  }
  if (m < e->lead_rings) return fail(ES_ERR_LIMIT, "lead rings exceed the strip limits");
  e->S = (uint32_t)S;
  e->R_outer = (uint64_t)(R / e->unit);
  e->strips = (uint64_t)(R / r0) / e->S;
  // Row-lane layout: the rings after the strip rings, while their radix product stays at
  // most 1024, are the sub rings a wave steps through (V steps per unit: short enough that a
  // window holds many units, so the last round of waves is short); the rest are the
  // superstrip rings.  It applies with every mask in LDS, a strip of at least 32 rows and at
  // least 16 steps, to one-word rows by default and to two-word rows (ring 0 of 33..64 slots)
  // on request (layout 2: which is faster depends on the wheel, so the caller times both).
  if (d->layout > 2) return fail(ES_ERR_INVALID, "layout must be 0, 1 or 2");
  uint32_t n_sub = 0;
  uint64_t V = 1;
  while (1 + m + n_sub < w && V * d->ring_radix[1 + m + n_sub] <= 1024) V *= d->ring_radix[1 + m + n_sub++];
  bool rows = (d->layout == 0 ? W == 1 : d->layout == 2 && W <= (uint32_t)kRowsMaxWords) &&
              d->global_masks == 0 && S >= 32 && V >= 16;
  if (d->layout == 2 && !rows)
    return fail(ES_ERR_LIMIT, "row-lane layout needs rows of at most two words, every mask in "
                              "LDS, a strip of at least 32 rows and at least 16 steps");
  e->rows = rows;
  e->V = rows ? (uint32_t)V : 1;
  e->Gl = rows ? (uint32_t)((S + 63) / 64) : 0;

  // ---- mask tables: rows [0, 2m) of 2W words (per ring-0 word: slots without / with an inner
  // wrap past M), one bit per sorted ring-0 slot.  Staged masks go to LDS, except that for
  // W = 1 the G smallest tables are read from global memory through L1.  A table of modulus m
  // keeps phase x at row P(x) = x * cinv mod m (and again at P + m), where cinv = c^-1 mod m
  // and c = -M mod m: subtracting M from n moves every table exactly one row on, so an outer
  // sum that wraps past M starts one row further on in every table (one shared offset), and
  // a row's second word is the next row's first.  A mask whose m is not prime to M has no
  // such c^-1 and is not staged (it is only a prefilter; its filter stays exact).
  const uint32_t budget_words = (d->lds_budget ? d->lds_budget : 65536u) / 4;
  // Carry layout.  Slots are sorted by e_i, so the slots of row h with h + e_i >= q0 are a
  // suffix [j(h), r0).  W = 1 and q0 <= 256: the table holds j(h) as one byte per h and the
  // kernel builds the carry word from it: the table spans at most 64 dwords, one per LDS
  // bank, so its random reads never conflict (7 % faster at k=399, q0 = 107, than word rows).
  // With q0 > 256 the bytes conflict too: the strip-lane kernel keeps word rows (the extra
  // instructions cost it 3 %), the row-lane kernel takes the bytes (a quarter of the banks'
  // load of word rows: the k=300..320 batch, 11 of whose 21 wheels have q0 = 343, 361 or 512,
  // ran 3 % faster).  Otherwise W <= 2: one row of W words per h (bit i set iff
  // j(h) <= i < r0).  W > 2: j(h) as u16, each carry word built from it (word rows would not
  // fit).  Row-lane rows of two words: j(h) <= 64 as one byte per h, each word's carry built
  // from it.
  auto carry_form = [&](bool rw) -> uint32_t {
    if (rw && W >= 2) return 1u;
    return (W == 1 && (q0 <= 256 || rw) && d->global_masks == 0) ? 1u : W <= 2 ? 0u : 2u;
  };
  uint32_t jbytes = carry_form(rows);
  e->kind = (jbytes == 1 ? 1u : 0u) | (e->lead > 1 ? 2u : 0u);
  if (e->lead > 1 && (W != 1 || d->global_masks != 0))
    return fail(ES_ERR_INVALID, "a lead block needs a ring 0 of at most 32 residues and "
                                "every mask in LDS");
  if (r0 > 0xffffu) return fail(ES_ERR_LIMIT, "ring-0 radix above 65535");
  uint32_t carry_words = jbytes ? (q0 * jbytes + 3) / 4 : q0 * W;
  std::vector<uint32_t> staged;  // indices into the descriptor's mask list
  {
    size_t used = carry_words + 8;
    const size_t cap_masks = rows ? (size_t)kRowsMaxMasks : (size_t)kMaxMasks;
    for (uint32_t f = 0; f < d->n_masks && staged.size() < cap_masks; ++f) {
      const size_t rows_words = rows ? (size_t)2 * W * (3 * (size_t)d->mask_mod[f] + 64) + 4
                                     : (size_t)4 * d->mask_mod[f] * W + 4;
      if (used + rows_words > budget_words) continue;  // stays exact-only
      if (inv_mod((uint32_t)((d->mask_mod[f] - mod_u128(M, d->mask_mod[f])) % d->mask_mod[f]),
                  d->mask_mod[f]) == 0)
        continue;                                      // m not prime to M
      used += rows_words;
      staged.push_back(f);
    }
  }
  if (rows && staged.empty()) {  // no mask tables: the strip-lane kernel
    if (d->layout == 2) return fail(ES_ERR_LIMIT, "row-lane layout needs at least one mask");
    rows = false;
    e->rows = false;
    e->V = 1;
    e->Gl = 0;
    jbytes = carry_form(false);
    e->kind = (jbytes == 1 ? 1u : 0u) | (e->lead > 1 ? 2u : 0u);
    carry_words = jbytes ? (q0 * jbytes + 3) / 4 : q0 * W;
  }
  const uint32_t nb = rows ? w - 1 - m - n_sub : w - 1 - m;
  const uint32_t first_base = rows ? 1 + m + n_sub : 1 + m;
  uint32_t G = d->global_masks;
  if (W != 1 || (G != 2 && G != 4) || staged.size() < G) G = 0;
  // the G smallest staged tables go global, kept in their relative order at the end
  std::vector<uint32_t> glob;
  if (G) {
    std::vector<uint32_t> bysize(staged);
    std::stable_sort(bysize.begin(), bysize.end(),
                     [&](uint32_t a, uint32_t b) { return d->mask_mod[a] < d->mask_mod[b]; });
    glob.assign(bysize.begin(), bysize.begin() + G);
  }
  std::vector<uint32_t> order_masks;
  for (uint32_t f : staged)
    if (std::find(glob.begin(), glob.end(), f) == glob.end()) order_masks.push_back(f);
  for (uint32_t f : staged)
    if (std::find(glob.begin(), glob.end(), f) != glob.end()) order_masks.push_back(f);
  // row-lane: an even modulus (a power of 2: at most one) goes last, with its own row map
  uint32_t last_even = 0;
  if (rows) {
    auto it = std::find_if(order_masks.begin(), order_masks.end(),
                           [&](uint32_t f) { return d->mask_mod[f] % 2 == 0; });
    if (it != order_masks.end()) {
      const uint32_t f = *it;
      order_masks.erase(it);
      order_masks.push_back(f);
      last_even = 1;
    }
  }
  std::vector<uint32_t> lds, gimg, mbase, mmod, msrc, minv;
  // Row-lane: the carry table opens the image, then block 0 of the tables (word 0 of each
  // mask, 16-byte aligned), then blocks 1..W-1 at the same offsets, blk_words apart.
  size_t blk_base = 0, blk_words = 0;
  std::vector<size_t> blk_off(order_masks.size(), 0);
  if (rows) {
    size_t o = ((size_t)carry_words + 3) & ~size_t(3);
    blk_base = o;
    for (size_t pos = 0; pos < order_masks.size(); ++pos) {
      blk_off[pos] = o - blk_base;
      o = (o + (size_t)2 * (3 * (size_t)d->mask_mod[order_masks[pos]] + 64) + 3) & ~size_t(3);
    }
    blk_words = o - blk_base;
    lds.assign(blk_base + (size_t)W * blk_words, 0);
  }
  for (size_t pos = 0; pos < order_masks.size(); ++pos) {
    const uint32_t f = order_masks[pos];
    const bool in_global = pos + G >= order_masks.size();
    std::vector<uint32_t>& img = in_global ? gimg : lds;
    const uint32_t mq = d->mask_mod[f], src = d->mask_src[f];
    const uint32_t corr = (mq - mod_u128(M, mq)) % mq;
    // row map: strip-lane keeps phase x at row x * (-M)^-1 (one row on per M subtracted), over
    // rows [0, 2m); row-lane at row x * lambda, lambda = 32 (-M)^-1 (32 rows on per M
    // subtracted: the same bank), over rows [0, 3m + 64)
    const uint32_t cinv = inv_mod(corr, mq);
    const uint32_t lam = rows && mq % 2 ? (uint32_t)((uint64_t)32 * cinv % mq) : cinv;
    const uint32_t xstep = inv_mod(lam, mq);            // phase of row 1
    const size_t n_rows = rows ? (size_t)3 * mq + 64 : (size_t)2 * mq;
    size_t off;
    if (rows) {
      off = blk_base + blk_off[pos];
    } else {
      off = (img.size() + 3) & ~size_t(3);  // 16-byte aligned rows
      img.resize(off + n_rows * 2 * W, 0);
    }
    const uint32_t* bm = d->mask_bits + d->mask_off[f];
    std::vector<uint32_t> t0m(r0);
    for (uint32_t i = 0; i < r0; ++i) t0m[i] = mod_u128(e->terms[order[i]], mq);
    for (uint32_t P = 0; P < mq; ++P) {
      const uint32_t x = (uint32_t)((uint64_t)P * xstep % mq);  // phase held at row P
      for (uint32_t i = 0; i < r0; ++i) {
        const uint32_t a = (uint32_t)(((uint64_t)x + t0m[i]) % mq);
        const uint32_t b = (uint32_t)(((uint64_t)a + corr) % mq);
        const uint32_t wi = i >> 5, bit = 1u << (i & 31);
        for (size_t r = P; r < n_rows; r += mq) {
          uint32_t* row = rows ? &img[off + wi * blk_words + 2 * r]
                               : &img[off + (size_t)2 * W * r + 2 * wi];
          if ((bm[a >> 5] >> (a & 31)) & 1u) row[0] |= bit;
          if ((bm[b >> 5] >> (b & 31)) & 1u) row[1] |= bit;
        }
      }
    }
    mbase.push_back((uint32_t)(off * 4));
    mmod.push_back(mq);
    msrc.push_back(src);
    minv.push_back(lam);
  }
  if (gimg.empty()) gimg.push_back(0);
  e->F = (uint32_t)mmod.size();
  e->G = G;
  e->mask_mod = mmod;
  e->mask_src = msrc;
  const uint32_t carry_off = rows ? 0u : (uint32_t)((lds.size() + 3) & ~size_t(3));
  if (!rows) lds.resize(carry_off + carry_words, 0);
  for (uint32_t h = 0; h < q0; ++h) {
    const uint32_t j = (uint32_t)(std::lower_bound(e_sorted.begin(), e_sorted.begin() + r0,
                                                   q0 - h) - e_sorted.begin());
    if (jbytes == 1) {
      lds[carry_off + h / 4] |= j << (8 * (h & 3));
    } else if (jbytes == 2) {
      lds[carry_off + h / 2] |= j << (16 * (h & 1));
    } else {
      for (uint32_t i = j; i < r0; ++i) lds[carry_off + (size_t)h * W + (i >> 5)] |= 1u << (i & 31);
    }
  }
  if (lds.size() > 40960) return fail(ES_ERR_LIMIT, "ring-0 tables exceed 160 KiB of LDS");
  // per workgroup: the tables, then each wave's candidate queue
  const uint32_t table_bytes = (uint32_t)((lds.size() * 4 + 15) & ~size_t(15));
  if ((uint64_t)e->S * W >= (1u << kQueueIndexBits))
    return fail(ES_ERR_LIMIT, "strip rows times ring-0 words exceed the queue index");

  // Every filter is decided: re-tested exactly, or covered in full by a staged mask (same
  // modulus and bitmap).  Filters left out of slow_order whose mask was not staged are
  // appended, so a short LDS can only cost speed.
  {
    std::vector<char> decided(d->n_filters, 0);
    for (uint32_t j : e->slow_order) decided[j] = 1;
    for (uint32_t u = 0; u < e->F; ++u) {
      const uint32_t src = msrc[u], mq = mmod[u];
      if (decided[src] || mq != d->filt_q[src]) continue;
      for (uint32_t f = 0; f < d->n_masks; ++f)
        if (d->mask_src[f] == src && d->mask_mod[f] == mq &&
            std::memcmp(d->mask_bits + d->mask_off[f], d->filt_bits + d->filt_off[src],
                        ((mq + 31) / 32) * 4) == 0)
          decided[src] = 1;
    }
    for (uint32_t j = 0; j < d->n_filters; ++j)
      if (!decided[j]) e->slow_order.push_back(j);
    e->n_slow = (uint32_t)e->slow_order.size();
  }

  // ---- base-ring terms in (h, l) form
  if (rows && nb > 64) return fail(ES_ERR_LIMIT, "more than 64 superstrip rings");
  std::vector<uint32_t> b_radix(std::max<uint32_t>(nb, 1), 0), b_off(std::max<uint32_t>(nb, 1), 0), t_h;
  std::vector<uint64_t> b_prefix(std::max<uint32_t>(nb, 1), 1);
  std::vector<uint64_t> t_l;
  for (uint32_t j = 0; j < nb; ++j) {
    const uint32_t ring = first_base + j;
    b_radix[j] = d->ring_radix[ring];
    b_off[j] = (uint32_t)t_h.size();
    if (j > 0) b_prefix[j] = b_prefix[j - 1] * b_radix[j - 1];
    for (uint32_t i = term_off[ring]; i < term_off[ring + 1]; ++i) {
      const u128 tv = e->terms[i];
      t_h.push_back((uint32_t)(tv / e->M0));
      const u128 lv = tv % e->M0;
      t_l.push_back((uint64_t)lv);
      t_l.push_back((uint64_t)(lv >> 64));
    }
  }
  if (t_h.empty()) {
    t_h.push_back(0);
    t_l.push_back(0);
    t_l.push_back(0);
  }

  // ---- strip rows over rings 1..m (ring 1 fastest): H, X, L (4 dwords), 0, 0, Q8_f.  X is
  // the idx correction of a lead block (0 otherwise): rings 1..j are then the reference's
  // rings 0..j-1 (lead = their radix product divides S), a slot of outer index o = (strip
  // base) + sg has idx r0 * (strip base) + sg % lead + lead * r0 * (sg / lead) + (slot
  // offset), and X = that - r0 * o - (slot offset) as a signed 32-bit value (|X| < r0 * S).
  const uint32_t F = e->F;
  const uint32_t stride = (8 + F + 7) & ~7u;
  std::vector<uint32_t> strip((size_t)S * stride, 0);
  std::vector<uint32_t> dig(m + 1, 0);
  for (uint64_t sg = 0; sg < S; ++sg) {
    u128 B = 0;
    for (uint32_t j = 0; j < m; ++j) {
      B += e->terms[term_off[1 + j] + dig[j]];
      if (B >= M) B -= M;
    }
    uint32_t* row = &strip[sg * stride];
    row[0] = (uint32_t)(B / e->M0);
    row[1] = (uint32_t)(int32_t)((int64_t)(sg % e->lead + (uint64_t)e->lead * r0 * (sg / e->lead)) -
                                 (int64_t)r0 * (int64_t)sg);
    const u128 L = B % e->M0;
    row[2] = (uint32_t)L;
    row[3] = (uint32_t)(L >> 32);
    row[4] = (uint32_t)(L >> 64);
    row[5] = (uint32_t)(L >> 96);
    for (uint32_t f = 0; f < F; ++f)  // row offset of phase B in table f (row-lane: word 0's)
      row[8 + f] = (rows ? 8u : 8u * W) *
                   (uint32_t)((uint64_t)mod_u128(B, mmod[f]) * minv[f] % mmod[f]);
    for (uint32_t j = 0; j < m; ++j) {  // odometer
      if (++dig[j] < d->ring_radix[1 + j]) break;
      dig[j] = 0;
    }
  }

  // ---- row-lane: step table over the sub rings (ring 1 + m fastest) and the lane groups
  std::vector<uint32_t> steps, perm;
  if (rows) {
    steps.assign((size_t)e->V * kStepDwords, 0);
    std::vector<uint32_t> sd(n_sub, 0);
    for (uint32_t v = 0; v < e->V; ++v) {
      u128 B = 0;
      for (uint32_t j = 0; j < n_sub; ++j) {
        B += e->terms[term_off[1 + m + j] + sd[j]];
        if (B >= M) B -= M;
      }
      uint32_t* st = &steps[(size_t)v * kStepDwords];
      st[0] = (uint32_t)(B / e->M0);
      const u128 z = e->M0 - B % e->M0;              // in (0, M0]
      st[1] = (uint32_t)z;
      st[2] = (uint32_t)(z >> 32);
      st[3] = (uint32_t)(z >> 64);
      st[4] = (uint32_t)(z >> 96);
      for (uint32_t f = 0; f < F; ++f)
        st[5 + f] = 8u * (uint32_t)((uint64_t)mod_u128(B, mmod[f]) * minv[f] % mmod[f]);
      for (uint32_t j = 0; j < n_sub; ++j) {  // odometer
        if (++sd[j] < d->ring_radix[1 + m + j]) break;
        sd[j] = 0;
      }
    }
    std::vector<uint32_t> bank((size_t)S * std::max<uint32_t>(F, 1), 0);
    for (uint64_t sg = 0; sg < S; ++sg)
      for (uint32_t f = 0; f < F; ++f) bank[sg * F + f] = (strip[sg * stride + 8 + f] / 8u) & 31u;
    perm = group_strip_rows(bank, (uint32_t)S, F, e->Gl);
  }

  // ---- mask records
  e->mask_info.assign((size_t)kMaskInfo * std::max<uint32_t>(F, 1), 0);
  for (uint32_t f = 0; f < F; ++f) {
    uint32_t* r = &e->mask_info[kMaskInfo * f];
    const uint32_t mq = mmod[f];
    r[0] = mq;
    r[1] = mbase[f];
    r[2] = mod_u128(e->M0, mq);
    r[3] = mod_u128((u128)1 << 32, mq);
    r[4] = mod_u128((u128)1 << 64, mq);
    r[5] = mod_u128((u128)1 << 96, mq);
    r[6] = minv[f];  // row of phase x: x * r[6] mod m
    r[7] = 0;
  }
  // ---- exact-filter records
  const uint32_t ns = e->n_slow;
  e->slow_info.assign((size_t)kSlowInfo * std::max<uint32_t>(ns, 1), 0);
  for (uint32_t j = 0; j < ns; ++j) {
    const uint32_t fi = e->slow_order[j], q = d->filt_q[fi];
    uint32_t* r = &e->slow_info[kSlowInfo * j];
    r[0] = q;
    r[1] = d->filt_off[fi];
    r[2] = mod_u128((u128)1 << 32, q);
    r[3] = mod_u128((u128)1 << 64, q);
    r[4] = mod_u128((u128)1 << 96, q);
    const double fd[4] = {(double)r[2], (double)r[3], (double)r[4], 1.0 / (double)q};
    std::memcpy(r + 8, fd, sizeof(fd));
  }
  bool slow_fp = true;
  for (uint32_t j = 0; j < ns; ++j) slow_fp = slow_fp && d->filt_q[e->slow_order[j]] < (1u << 18);
  std::vector<uint32_t> fbits(d->filt_bits, d->filt_bits + d->n_filt_words);
  if (fbits.empty()) fbits.push_back(0);

  // ---- device upload: one blob
  Blob bl;
  const size_t o_lds = bl.put(lds.data(), lds.size() * 4);
  const size_t o_gm = bl.put(gimg.data(), gimg.size() * 4);
  const size_t o_strip = bl.put(strip.data(), strip.size() * 4);
  const size_t o_brad = bl.put(b_radix.data(), b_radix.size() * 4);
  const size_t o_boff = bl.put(b_off.data(), b_off.size() * 4);
  const size_t o_th = bl.put(t_h.data(), t_h.size() * 4);
  const size_t o_tl = bl.put(t_l.data(), t_l.size() * 8);
  const size_t o_mask = bl.put(e->mask_info.data(), e->mask_info.size() * 4);
  const size_t o_e0 = bl.put(e_sorted.data(), e_sorted.size() * 4);
  const size_t o_ids = bl.put(inner_ids.data(), inner_ids.size() * 4);
  const size_t o_slow = bl.put(e->slow_info.data(), e->slow_info.size() * 4);
  const size_t o_sbits = bl.put(fbits.data(), fbits.size() * 4);
  const size_t o_bpre = bl.put(b_prefix.data(), b_prefix.size() * 8);
  const size_t o_rp = bl.put(nullptr, sizeof(RowPlan));
  const size_t o_steps = bl.put(steps.data(), steps.size() * 4);
  const size_t o_perm = bl.put(perm.data(), perm.size() * 4);
  const size_t o_cold = bl.put(nullptr, sizeof(ColdPlan));
  const size_t o_cnt = bl.put(nullptr, 24);

  ES_HIP(hipGetDevice(&e->device));
  hipDeviceProp_t prop;
  ES_HIP(hipGetDeviceProperties(&prop, e->device));
  if (rows) {
    e->rows_kern = pick_rows((int)e->F, e->kind, W, false);
    if (!e->rows_kern) return fail(ES_ERR_LIMIT, "no row-lane kernel for this mask count");
    e->kern = nullptr;
  } else {
    e->kern = pick_kernel((int)e->F, W, G, e->kind);
    if (!e->kern) return fail(ES_ERR_LIMIT, "no kernel for this mask count");
  }
  {
    // The workgroup size (unless given) and, for one-word row-lane rows, the kernel instance
    // (registers for 6 or 8 waves per SIMD) that keep the most waves resident per CU; ties go
    // to the 6-wave instance (no spills), then to the smaller workgroup.  Every workgroup
    // holds its own copy of the tables, so large tables favour larger workgroups; registers
    // and the CU's wave slots bound the rest.
    struct Cand { const void* fn; bool wave8; };
    std::vector<Cand> cands{{rows ? (const void*)e->rows_kern : (const void*)e->kern, false}};
    if (rows && W == 1) cands.push_back({(const void*)pick_rows((int)e->F, e->kind, W, true), true});
    const std::vector<uint32_t> sizes = e->threads ? std::vector<uint32_t>{e->threads}
                                                   : std::vector<uint32_t>{256u, 512u, 1024u};
    uint32_t best_waves = 0, best_t = 0, best6_waves = 0, best6_t = 0;
    bool best_w8 = false;
    for (const Cand& c : cands) {
      if (!c.fn) continue;
      for (uint32_t t : sizes) {
        const uint32_t lds_t = table_bytes + (t / 64) * kWaveQueueBytes;
        if (lds_t > prop.sharedMemPerBlock) continue;
        if (lds_t > 65536)
          ES_HIP(hipFuncSetAttribute(c.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)prop.sharedMemPerBlock));
        int nb = 0;
        ES_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, c.fn, (int)t, lds_t));
        const uint32_t waves = (uint32_t)nb * (t / 64);
        if (waves > best_waves) {
          best_waves = waves;
          best_t = t;
          best_w8 = c.wave8;
        }
        if (!c.wave8 && waves > best6_waves) {
          best6_waves = waves;
          best6_t = t;
        }
      }
    }
    const uint32_t given = e->threads;
    if (!e->threads) e->threads = best_t ? best_t : 256;  // nothing fits: reported below
    if (best_w8) {
      e->wave8 = true;
      e->rows_kern6 = e->rows_kern;
      e->threads6 = given ? given : (best6_t ? best6_t : e->threads);
      e->lds_bytes6 = table_bytes + (e->threads6 / 64) * kWaveQueueBytes;
      e->rows_kern = pick_rows((int)e->F, e->kind, W, true);
    }
  }
  const void* kfn = rows ? (const void*)e->rows_kern : (const void*)e->kern;
  e->lds_bytes = table_bytes + (e->threads / 64) * kWaveQueueBytes;
  if (e->lds_bytes > prop.sharedMemPerBlock)
    return fail(ES_ERR_LIMIT, "LDS tables and candidate queues exceed the device's "
                              "per-workgroup LDS (fewer threads or a smaller budget)");
  e->max_grid = 1u << 30;  // one strip per thread; the dispatcher balances the tail
  ES_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  ES_HIP(hipMalloc(&e->dblob, bl.bytes.size()));
  if (e->out_cap > (1ull << 32)) return fail(ES_ERR_LIMIT, "out_cap above 2^32 survivors");
  ES_HIP(hipMalloc(&e->d_out, std::max<uint64_t>(e->out_cap, 1) * 24));
  uint8_t* base = (uint8_t*)e->dblob;
  ColdPlan cold{};
  cold.M0_lo = (uint64_t)e->M0;
  cold.M0_hi = (uint64_t)(e->M0 >> 64);
  cold.q0 = q0;
  cold.r0 = r0;
  cold.n_slow = ns;
  cold.stride = stride;
  cold.slow_fp = slow_fp ? 1u : 0u;
  cold.strip = (const uint32_t*)(base + o_strip);
  cold.e0 = (const uint32_t*)(base + o_e0);
  cold.inner_ids = (const uint32_t*)(base + o_ids);
  cold.slow = (const uint32_t*)(base + o_slow);
  cold.slow_bits = (const uint32_t*)(base + o_sbits);
  cold.out = e->d_out;
  cold.out_n = (unsigned long long*)(base + o_cnt);
  cold.out_s = e->d_out + std::max<uint64_t>(e->out_cap, 1);
  cold.out_cap = e->out_cap;
  e->cold = cold;
  std::memcpy(bl.bytes.data() + o_cold, &cold, sizeof(cold));
  ES_HIP(hipMemcpy(e->dblob, bl.bytes.data(), bl.bytes.size(), hipMemcpyHostToDevice));
  e->d_count = (unsigned long long*)(base + o_cnt);
  e->d_mask_info = (uint32_t*)(base + o_mask);
  e->d_slow_info = (uint32_t*)(base + o_slow);

  HotPlan& P = e->hp;
  P.M0_lo = (uint64_t)e->M0;
  P.M0_hi = (uint64_t)(e->M0 >> 64);
  P.q0 = q0;
  P.r0 = r0;
  P.S = e->S;
  P.nb = nb;
  P.W = W;
  P.stride = stride;
  P.lds_words = (uint32_t)lds.size();
  P.carry_base = carry_off * 4;
  P.lds_src = (const uint32_t*)(base + o_lds);
  P.strip = (const uint32_t*)(base + o_strip);
  P.base_radix = (const uint32_t*)(base + o_brad);
  P.base_off = (const uint32_t*)(base + o_boff);
  P.term_h = (const uint32_t*)(base + o_th);
  P.term_l = (const uint64_t*)(base + o_tl);
  P.masks = (const uint32_t*)(base + o_mask);
  P.gmask = (const uint32_t*)(base + o_gm);
  P.cold = (const ColdPlan*)(base + o_cold);
  if (e->lds_bytes > 65536)  // the device's limit, so engines sharing a kernel never lower it
    ES_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)prop.sharedMemPerBlock));
  if (rows) {
    RowPlan& Q = e->rp;
    Q.M0_lo = P.M0_lo;
    Q.M0_hi = P.M0_hi;
    Q.q0 = q0;
    Q.r0 = r0;
    Q.S = e->S;
    Q.G = e->Gl;
    Q.V = e->V;
    Q.nb = nb;
    Q.lds_words = P.lds_words;
    Q.carry_base = P.carry_base;
    Q.stride = stride;
    Q.last_even = last_even;
    Q.word_bytes = (uint32_t)(blk_words * 4);
    Q.lds_src = P.lds_src;
    Q.perm = (const uint32_t*)(base + o_perm);
    Q.strip = P.strip;
    Q.steps = (const uint32_t*)(base + o_steps);
    Q.base_radix = P.base_radix;
    Q.base_off = P.base_off;
    Q.base_prefix = (const uint64_t*)(base + o_bpre);
    Q.term_h = P.term_h;
    Q.term_l = P.term_l;
    Q.masks = P.masks;
    Q.cold = P.cold;
    e->d_rp = (RowPlan*)(base + o_rp);
    ES_HIP(hipMemcpy(e->d_rp, &Q, sizeof(RowPlan), hipMemcpyHostToDevice));
  }
  e->launch_outer = std::max<uint64_t>(1, ((uint64_t)1 << (d->launch_log2 ? d->launch_log2 : 37)) / r0);
  e->ev_pool.resize(512);
  for (auto& ev : e->ev_pool) ES_HIP(hipEventCreate(&ev));
  ES_HIP(hipMemsetAsync(e->d_count, 0, 24, e->stream));
  std::vector<uint32_t> tc0(std::max<uint32_t>(d->n_filters, 1), 0);
  const int rc = upload_tconst(e, tc0);
  if (rc) return rc;
  *out = guard.release();
  return ES_OK;
}

int es_engine_destroy(es_engine* e) {
  if (!e) return ES_OK;
  if (e->stream) hipStreamSynchronize(e->stream);
  for (auto& ev : e->ev_pool) hipEventDestroy(ev);
  for (auto& sl : e->slot)
    if (sl.done) hipEventDestroy(sl.done);
  if (e->slot_mem) hipFree(e->slot_mem);
  if (e->h_hdr) hipHostFree(e->h_hdr);
  if (e->copy_stream) hipStreamDestroy(e->copy_stream);
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
  info->global_masks = e->G;
  info->layout = e->rows ? 2 : 1;
  info->steps = e->rows ? e->V : 0;
  info->groups = e->rows ? e->Gl : 0;
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
  e->t_known = false;
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
  e->t_known = true;
  return upload_tconst(e, tc);
}

int es_engine_reset(es_engine* e) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  e->count_valid = false;
  ES_HIP(hipMemsetAsync(e->d_count, 0, 16, e->stream));  // count and minimum, not the floor
  return ES_OK;
}

int es_engine_scan_window(es_engine* e, uint64_t idx_lo, uint64_t idx_hi) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  if (idx_hi > (uint64_t)e->R || idx_lo > idx_hi) return fail(ES_ERR_INVALID, "window outside [0, R)");
  return scan_range(e, idx_lo, idx_hi, nullptr, nullptr);
}

int es_engine_scan_windows(es_engine* e, const uint64_t* bounds, uint32_t n_windows,
                           uint64_t* ids, uint64_t cap, uint64_t* counts, uint64_t* min_s,
                           uint8_t* has_min) {
  if (!e || (n_windows && (!bounds || !counts)) || (cap && !ids))
    return fail(ES_ERR_INVALID, "null argument");
  if ((min_s || has_min) && !e->t_known)
    return fail(ES_ERR_INVALID, "the minimum needs the block index: set it with "
                                "es_engine_set_block (es_engine_set_tconst leaves it unknown)");
  for (uint32_t i = 0; i < n_windows; ++i)
    if (bounds[2 * i + 1] > (uint64_t)e->R || bounds[2 * i] > bounds[2 * i + 1])
      return fail(ES_ERR_INVALID, "window outside [0, R)");
  if (n_windows == 0) return ES_OK;
  int rc = setup_slots(e);
  if (rc) return rc;
  // both slots start from zero with the block's floor (the least s that counts for the minimum)
  uint64_t* init = e->h_hdr + 8;
  init[0] = 0;
  init[1] = 0;
  init[2] = e->floor;
  init[3] = 0;
  for (auto& sl : e->slot)
    ES_HIP(hipMemcpyAsync(sl.d_ctr, init, 32, hipMemcpyHostToDevice, e->stream));
  ES_HIP(hipStreamSynchronize(e->stream));
  auto submit = [&](uint32_t i) -> int {
    es_engine::Slot& sl = e->slot[i & 1];
    const int r = scan_range(e, bounds[2 * i], bounds[2 * i + 1], sl.d_cold, sl.d_rp);
    if (r) return r;
    hipLaunchKernelGGL(es_slot_finish_kernel, dim3(1), dim3(64), 0, e->stream, sl.d_ctr,
                       (const unsigned long long*)(sl.d_out + std::max<uint64_t>(e->out_cap, 1)),
                       e->h_hdr_dev + 4 * (i & 1));
    ES_HIP(hipGetLastError());
    ES_HIP(hipEventRecord(sl.done, e->stream));
    return ES_OK;
  };
  // window i's results, read while window i + 1 runs; errors are kept until the stream drains
  uint64_t used = 0;
  int err = ES_OK;
  auto finish = [&](uint32_t i) -> int {
    es_engine::Slot& sl = e->slot[i & 1];
    ES_HIP(hipEventSynchronize(sl.done));
    const uint64_t* h = e->h_hdr + 4 * (i & 1);
    const uint64_t c = h[0];
    counts[i] = c;
    if (min_s) {
      min_s[2 * i] = h[1];
      min_s[2 * i + 1] = h[2];
    }
    if (has_min) has_min[i] = (uint8_t)h[3];
    if (c > e->out_cap || c > cap - std::min(cap, used)) {
      if (!err) err = fail(ES_ERR_OVERFLOW, "survivor buffer overflow");
      return ES_OK;
    }
    if (c && !err) {
      ES_HIP(hipMemcpyAsync(ids + used, sl.d_out, c * 8, hipMemcpyDeviceToHost, e->copy_stream));
      ES_HIP(hipStreamSynchronize(e->copy_stream));
    }
    used += c;
    return ES_OK;
  };
  rc = submit(0);
  for (uint32_t i = 1; i < n_windows && !rc; ++i) {
    rc = submit(i);
    if (!rc) rc = finish(i - 1);
  }
  if (rc) {  // a launch failed: let the windows already queued finish before returning
    hipStreamSynchronize(e->stream);
    return rc;
  }
  rc = finish(n_windows - 1);
  return rc ? rc : err;
}

int es_engine_scan(es_engine* e, uint64_t ostart, uint64_t ocount) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  if (ostart > e->R_outer || ocount > e->R_outer - ostart)
    return fail(ES_ERR_INVALID, "outer range outside [0, R_outer)");
  return es_engine_scan_window(e, ostart * e->unit, (ostart + ocount) * e->unit);
}

int es_engine_sync(es_engine* e) {
  if (!e) return fail(ES_ERR_INVALID, "null engine");
  ES_HIP(hipStreamSynchronize(e->stream));
  return ES_OK;
}

int es_engine_count(es_engine* e, uint64_t* n) {
  if (!e || !n) return fail(ES_ERR_INVALID, "null argument");
  if (e->count_valid) {  // nothing launched or reset since the last read
    *n = e->last_count;
    return ES_OK;
  }
  unsigned long long c = 0;
  ES_HIP(hipMemcpyAsync(&c, e->d_count, 8, hipMemcpyDeviceToHost, e->stream));
  ES_HIP(hipStreamSynchronize(e->stream));
  *n = c;
  e->last_count = c;
  e->count_valid = true;
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
  if (has_min) *has_min = 0;
  if ((min_s || has_min) && !e->t_known)
    return fail(ES_ERR_INVALID, "the minimum needs the block index: set it with "
                                "es_engine_set_block (es_engine_set_tconst leaves it unknown)");
  // the counter and the minimum's slot, read together after the stream drains
  unsigned long long ctr[2] = {0, 0};
  ES_HIP(hipMemcpyAsync(ctr, e->d_count, 16, hipMemcpyDeviceToHost, e->stream));
  ES_HIP(hipStreamSynchronize(e->stream));
  e->last_count = ctr[0];
  e->count_valid = true;
  const uint64_t c = ctr[0];
  *n = c;
  if (c > e->out_cap) return fail(ES_ERR_OVERFLOW, "survivor buffer overflow");
  if (ids && c && cap) {
    const int rc = es_engine_read(e, 0, std::min<uint64_t>(c, cap), ids);
    if (rc) return rc;
  }
  if (ctr[1] && (min_s || has_min)) {
    // the record of the least residue s (n = t*M + s > k + 1), reduced on the device
    uint64_t rec[2] = {0, 0};
    ES_HIP(hipMemcpyAsync(rec, e->cold.out_s + 2 * (ctr[1] - 1), 16, hipMemcpyDeviceToHost,
                          e->stream));
    ES_HIP(hipStreamSynchronize(e->stream));
    if (min_s) {
      min_s[0] = rec[0];
      min_s[1] = rec[1];
    }
    if (has_min) *has_min = 1;
  }
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
  if (ms) *ms = tot;
  if (launches) *launches = e->ev_used / 2;
  if (positions) *positions = e->positions;
  e->ev_used = 0;
  e->launches = 0;
  e->positions = 0;
  return ES_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------ batches

struct es_batch {
  struct Job {
    es_engine* e = nullptr;
    uint64_t t = 0, idx_lo = 0, idx_hi = 0, cap = 0;
    void* dmem = nullptr;                   // mask records, slow records, cold plan, counter
    size_t o_mi = 0, o_si = 0;              // offsets of the mask and slow records in dmem
    unsigned long long* d_out = nullptr;
    unsigned long long* d_count = nullptr;
    HotPlan hp{};
    RowPlan rp{};                           // row-lane engines
  };
  hipStream_t stream = nullptr;
  // groups of different shapes are independent: they run concurrently on side streams
  // (default priority), forked from and joined back to `stream` with events
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[kSide] = {nullptr, nullptr, nullptr};
  uint64_t out_cap = 1u << 16;
  std::vector<Job> jobs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* d_tables = nullptr;  // plan views and job table of the last run (grown as needed)
  size_t tables_bytes = 0;
  size_t tables_hot_bytes = 0, tables_plan_bytes = 0, tables_job_bytes = 0;
  hipGraph_t graph = nullptr;          // captured launches (es_batch_capture)
  hipGraphExec_t graph_exec = nullptr;
  void drop_graph() {
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    if (graph) hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }
};

// Zero the survivor counters of a captured batch (a kernel node: graphs here hold no memset
// nodes).
__global__ void es_zero_counters_kernel(unsigned long long* const* __restrict__ ptrs, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    ptrs[i][0] = 0ull;  // survivors
    ptrs[i][1] = 0ull;  // slot of the minimum
  }
}

// Block constants of a job: the mask and slow-filter records of engine e with t*M mod q.
static uint64_t job_floor(const es_engine* e, uint64_t t) { return t == 0 ? (uint64_t)e->k + 2 : 0; }
static void job_records(const es_engine* e, uint64_t t, std::vector<uint32_t>& mi,
                        std::vector<uint32_t>& si) {
  std::vector<uint32_t> tc(std::max<uint32_t>(e->n_filters, 1), 0);
  for (uint32_t f = 0; f < e->n_filters; ++f) {
    const uint32_t q = e->filt_q[f];
    tc[f] = (uint32_t)(((uint64_t)(t % q) * mod_u128(e->M, q)) % q);
  }
  mi = e->mask_info;
  si = e->slow_info;
  for (uint32_t f = 0; f < e->F; ++f) mi[kMaskInfo * f + 7] = tc[e->mask_src[f]] % e->mask_mod[f];
  for (uint32_t s = 0; s < e->n_slow; ++s) si[kSlowInfo * s + 5] = tc[e->slow_order[s]];
}

extern "C" {

int es_batch_create(uint64_t out_cap_per_job, es_batch** out) {
  if (!out) return fail(ES_ERR_INVALID, "null argument");
  *out = nullptr;
  std::unique_ptr<es_batch> b(new es_batch());
  b->out_cap = out_cap_per_job ? out_cap_per_job : (1u << 16);
  ES_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  for (int i = 0; i < es_batch::kSide; ++i) {
    ES_HIP(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking));
    ES_HIP(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming));
  }
  ES_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  ES_HIP(hipEventCreate(&b->ev0));
  ES_HIP(hipEventCreate(&b->ev1));
  *out = b.release();
  return ES_OK;
}

int es_batch_destroy(es_batch* b) {
  if (!b) return ES_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  for (auto& j : b->jobs) {
    hipFree(j.dmem);
    hipFree(j.d_out);
  }
  hipFree(b->d_tables);
  b->drop_graph();
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  for (int i = 0; i < es_batch::kSide; ++i) {
    if (b->side[i]) {
      hipStreamSynchronize(b->side[i]);
      hipStreamDestroy(b->side[i]);
    }
    if (b->join[i]) hipEventDestroy(b->join[i]);
  }
  if (b->fork) hipEventDestroy(b->fork);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return ES_OK;
}

int es_batch_add(es_batch* b, es_engine* e, uint64_t t, uint64_t idx_lo, uint64_t idx_hi,
                 uint32_t* job_id) {
  if (!b || !e) return fail(ES_ERR_INVALID, "null argument");
  if (idx_hi > (uint64_t)e->R || idx_lo >= idx_hi) return fail(ES_ERR_INVALID, "window outside [0, R)");
  if (e->G != 0) return fail(ES_ERR_LIMIT, "batched launches read every mask from LDS (G = 0)");
  es_batch::Job j;
  j.e = e;
  j.t = t;
  j.idx_lo = idx_lo;
  j.idx_hi = idx_hi;
  j.cap = b->out_cap;
  std::vector<uint32_t> mi, si;
  job_records(e, t, mi, si);
  Blob bl;
  const size_t o_mi = bl.put(mi.data(), mi.size() * 4);
  const size_t o_si = bl.put(si.data(), si.size() * 4);
  j.o_mi = o_mi;
  j.o_si = o_si;
  const size_t o_cold = bl.put(nullptr, sizeof(ColdPlan));
  const size_t o_cnt = bl.put(nullptr, 24);
  const uint64_t fl = job_floor(e, t);
  std::memcpy(bl.bytes.data() + o_cnt + 16, &fl, 8);
  ES_HIP(hipMalloc(&j.dmem, bl.bytes.size()));
  ES_HIP(hipMalloc((void**)&j.d_out, j.cap * 24));
  uint8_t* base = (uint8_t*)j.dmem;
  ColdPlan cold = e->cold;
  cold.slow = (const uint32_t*)(base + o_si);
  cold.out = j.d_out;
  cold.out_n = (unsigned long long*)(base + o_cnt);
  cold.out_s = j.d_out + j.cap;
  cold.out_cap = j.cap;
  std::memcpy(bl.bytes.data() + o_cold, &cold, sizeof(cold));
  ES_HIP(hipMemcpy(j.dmem, bl.bytes.data(), bl.bytes.size(), hipMemcpyHostToDevice));
  j.d_count = (unsigned long long*)(base + o_cnt);
  j.hp = e->hp;
  j.hp.masks = (const uint32_t*)(base + o_mi);
  j.hp.cold = (const ColdPlan*)(base + o_cold);
  j.rp = e->rp;
  j.rp.masks = j.hp.masks;
  j.rp.cold = j.hp.cold;
  if (job_id) *job_id = (uint32_t)b->jobs.size();
  b->jobs.push_back(j);
  b->drop_graph();  // a captured graph no longer covers every job
  return ES_OK;
}

int es_batch_set_blocks(es_batch* b, const uint64_t* t, uint32_t n_jobs) {
  if (!b || (n_jobs && !t) || n_jobs != b->jobs.size())
    return fail(ES_ERR_INVALID, "one block per job expected");
  // all records staged in one host buffer, copied in order, one synchronisation
  std::vector<std::vector<uint32_t>> stage(2 * (size_t)n_jobs);
  std::vector<uint64_t> floors(n_jobs);
  for (uint32_t i = 0; i < n_jobs; ++i) {
    es_batch::Job& j = b->jobs[i];
    job_records(j.e, t[i], stage[2 * i], stage[2 * i + 1]);
    floors[i] = job_floor(j.e, t[i]);
    uint8_t* base = (uint8_t*)j.dmem;
    ES_HIP(hipMemcpyAsync(base + j.o_mi, stage[2 * i].data(), stage[2 * i].size() * 4,
                          hipMemcpyHostToDevice, b->stream));
    ES_HIP(hipMemcpyAsync(base + j.o_si, stage[2 * i + 1].data(), stage[2 * i + 1].size() * 4,
                          hipMemcpyHostToDevice, b->stream));
    ES_HIP(hipMemcpyAsync(j.d_count + 2, &floors[i], 8, hipMemcpyHostToDevice, b->stream));
  }
  ES_HIP(hipStreamSynchronize(b->stream));
  for (uint32_t i = 0; i < n_jobs; ++i) b->jobs[i].t = t[i];
  return ES_OK;
}

int es_batch_set_block(es_batch* b, uint32_t job, uint64_t t) {
  if (!b || job >= b->jobs.size()) return fail(ES_ERR_INVALID, "bad job");
  es_batch::Job& j = b->jobs[job];
  std::vector<uint32_t> mi, si;
  job_records(j.e, t, mi, si);
  uint8_t* base = (uint8_t*)j.dmem;
  const uint64_t fl = job_floor(j.e, t);
  // ordered after earlier work on the batch's stream (replays read these records)
  ES_HIP(hipMemcpyAsync(base + j.o_mi, mi.data(), mi.size() * 4, hipMemcpyHostToDevice, b->stream));
  ES_HIP(hipMemcpyAsync(base + j.o_si, si.data(), si.size() * 4, hipMemcpyHostToDevice, b->stream));
  ES_HIP(hipMemcpyAsync(j.d_count + 2, &fl, 8, hipMemcpyHostToDevice, b->stream));
  ES_HIP(hipStreamSynchronize(b->stream));
  j.t = t;
  return ES_OK;
}

namespace {

// Launch groups of a batch with their device tables uploaded (prepare_batch), and the launch
// of every group on the batch's stream (launch_batch: no allocation or host synchronisation,
// so it can be captured into a graph).
struct BatchLaunch {
  const void* fn;
  bool rows;
  uint32_t threads, plan_off, job_off, n, lds_max;
  uint64_t wgs;
};

int prepare_batch(es_batch* b, std::vector<BatchLaunch>& out) {
  out.clear();
  // Group jobs by kernel instance, workgroup size and occupancy class: a launch allocates its
  // largest plan's LDS for every workgroup, so only plans that fit as many workgroups per CU
  // (2048 lanes and 160 KiB of LDS per CU) share one; a small plan is never held to a large
  // plan's occupancy.
  std::vector<std::vector<uint32_t>> groups;
  struct Key { const void* fn; uint32_t threads, occ; };
  std::vector<Key> keys;
  for (uint32_t i = 0; i < b->jobs.size(); ++i) {
    const es_engine* e = b->jobs[i].e;
    const void* fn = e->rows ? (const void*)pick_rows_batch((int)e->F, e->kind, e->W, e->wave8)
                             : (const void*)pick_grid_batch((int)e->F, e->W, 0, e->kind);
    if (!fn) return fail(ES_ERR_LIMIT, "no batched kernel for this mask count");
    const uint32_t occ = std::min<uint32_t>(2048u / e->threads,
                                            163840u / std::max<uint32_t>(e->lds_bytes, 1));
    size_t g = 0;
    while (g < keys.size() &&
           !(keys[g].fn == fn && keys[g].threads == e->threads && keys[g].occ == occ))
      ++g;
    if (g == keys.size()) {
      keys.push_back({fn, e->threads, occ});
      groups.emplace_back();
    }
    groups[g].push_back(i);
  }
  // per group: plan views [plan_off, +n) (strip-lane groups in the HotPlan array, row-lane
  // groups in the RowPlan array) and job entries [job_off, +n) in one device table, then the
  // jobs' counter addresses
  struct Launch { bool rows; uint32_t plan_off, job_off, n, lds_max; uint64_t wgs; };
  std::vector<Launch> runs;
  std::vector<HotPlan> plans;
  std::vector<RowPlan> rplans;
  std::vector<GridJob> gj;
  for (size_t g = 0; g < groups.size(); ++g) {
    // every job's strips (strip-lane) or units (row-lane: one per wave) as consecutive
    // workgroups of one launch
    const uint32_t threads = keys[g].threads;
    const bool rows = b->jobs[groups[g][0]].e->rows;
    Launch L{rows, (uint32_t)(rows ? rplans.size() : plans.size()), (uint32_t)gj.size(),
             (uint32_t)groups[g].size(), 4, 0};
    uint64_t wgs = 0;
    for (uint32_t ji : groups[g]) {
      const es_batch::Job& j = b->jobs[ji];
      const es_engine* e = j.e;
      GridJob J{};
      if (rows) {
        J.plan = (uint32_t)(rplans.size() - L.plan_off);
        rplans.push_back(j.rp);
      } else {
        J.plan = (uint32_t)(plans.size() - L.plan_off);
        plans.push_back(j.hp);
      }
      L.lds_max = std::max(L.lds_max, e->lds_bytes);
      outer_range(e, j.idx_lo, j.idx_hi, &J.o_lo, &J.o_hi);
      J.idx_lo = j.idx_lo;
      J.idx_hi = j.idx_hi;
      J.wg_start = wgs;
      if (rows) {
        const uint64_t span = (uint64_t)e->S * e->V;
        J.u_lo = J.o_lo / span;
        J.u_hi = (J.o_hi + span - 1) / span;
        wgs += ((uint64_t)e->Gl * (J.u_hi - J.u_lo) + threads / 64 - 1) / (threads / 64);
      } else {
        J.u_lo = J.o_lo / e->S;
        J.u_hi = (J.o_hi + e->S - 1) / e->S;
        wgs += (J.u_hi - J.u_lo + threads - 1) / threads;
      }
      gj.push_back(J);
    }
    if (wgs > 0x7fffffffull) return fail(ES_ERR_LIMIT, "batch too large for one grid");
    L.wgs = wgs;
    runs.push_back(L);
  }
  const size_t hot_bytes = (plans.size() * sizeof(HotPlan) + 255) & ~size_t(255);
  const size_t row_bytes = (rplans.size() * sizeof(RowPlan) + 255) & ~size_t(255);
  const size_t plan_bytes = hot_bytes + row_bytes;
  const size_t job_bytes = (gj.size() * sizeof(GridJob) + 255) & ~size_t(255);
  std::vector<unsigned long long*> ctrs;
  for (auto& j : b->jobs) ctrs.push_back(j.d_count);
  const size_t need = plan_bytes + job_bytes + std::max<size_t>(ctrs.size(), 1) * 8;
  if (need > b->tables_bytes) {
    b->drop_graph();
    if (b->d_tables) ES_HIP(hipFree(b->d_tables));
    b->d_tables = nullptr;
    ES_HIP(hipMalloc(&b->d_tables, need));
    b->tables_bytes = need;
  }
  uint8_t* tb = (uint8_t*)b->d_tables;
  if (!gj.empty()) {
    if (!plans.empty())
      ES_HIP(hipMemcpyAsync(tb, plans.data(), plans.size() * sizeof(HotPlan),
                            hipMemcpyHostToDevice, b->stream));
    if (!rplans.empty())
      ES_HIP(hipMemcpyAsync(tb + hot_bytes, rplans.data(), rplans.size() * sizeof(RowPlan),
                            hipMemcpyHostToDevice, b->stream));
    ES_HIP(hipMemcpyAsync(tb + plan_bytes, gj.data(), gj.size() * sizeof(GridJob),
                          hipMemcpyHostToDevice, b->stream));
    ES_HIP(hipMemcpyAsync(tb + plan_bytes + job_bytes, ctrs.data(), ctrs.size() * 8,
                          hipMemcpyHostToDevice, b->stream));
  }
  ES_HIP(hipStreamSynchronize(b->stream));
  for (size_t g = 0; g < runs.size(); ++g) {
    const Launch& L = runs[g];
    if (L.lds_max > 65536) {  // the device's limit, so batches sharing a kernel never lower it
      int dev = 0, lim = 0;
      ES_HIP(hipGetDevice(&dev));
      ES_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
      ES_HIP(hipFuncSetAttribute(keys[g].fn, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
    }
    out.push_back({keys[g].fn, L.rows, keys[g].threads, L.plan_off, L.job_off, L.n, L.lds_max,
                   L.wgs});
  }
  b->tables_hot_bytes = hot_bytes;
  b->tables_plan_bytes = plan_bytes;
  b->tables_job_bytes = job_bytes;
  return ES_OK;
}

int launch_batch(es_batch* b, const std::vector<BatchLaunch>& runs, bool zero_kernel) {
  uint8_t* tb = (uint8_t*)b->d_tables;
  const uint32_t nj = (uint32_t)b->jobs.size();
  if (zero_kernel && nj) {
    unsigned long long* const* pp =
        (unsigned long long* const*)(tb + b->tables_plan_bytes + b->tables_job_bytes);
    hipLaunchKernelGGL(es_zero_counters_kernel, dim3((nj + 255) / 256), dim3(256), 0, b->stream,
                       pp, nj);
    ES_HIP(hipGetLastError());
  }
  // group g runs on stream g mod (kSide + 1): 0 = the batch's stream, 1.. = side streams
  const int ns = (int)std::min<size_t>(runs.size(), es_batch::kSide + 1);
  if (ns > 1) {
    ES_HIP(hipEventRecord(b->fork, b->stream));
    for (int i = 0; i + 1 < ns; ++i) ES_HIP(hipStreamWaitEvent(b->side[i], b->fork, 0));
  }
  for (size_t g = 0; g < runs.size(); ++g) {
    const BatchLaunch& L = runs[g];
    const int si = (int)(g % (size_t)ns);
    hipStream_t st = si == 0 ? b->stream : b->side[si - 1];
    const void* pp = L.rows ? (const void*)((const RowPlan*)(tb + b->tables_hot_bytes) + L.plan_off)
                            : (const void*)((const HotPlan*)tb + L.plan_off);
    const GridJob* jp = (const GridJob*)(tb + b->tables_plan_bytes) + L.job_off;
    uint32_t n = L.n;
    void* args[] = {&pp, &jp, &n};
    ES_HIP(hipLaunchKernel((const void*)L.fn, dim3((uint32_t)std::max<uint64_t>(L.wgs, 1)),
                           dim3(L.threads), args, L.lds_max, st));
    ES_HIP(hipGetLastError());
  }
  for (int i = 0; i + 1 < ns; ++i) {
    ES_HIP(hipEventRecord(b->join[i], b->side[i]));
    ES_HIP(hipStreamWaitEvent(b->stream, b->join[i], 0));
  }
  return ES_OK;
}

}  // namespace

int es_batch_run(es_batch* b, double* ms, uint32_t* launches) {
  if (!b) return fail(ES_ERR_INVALID, "null argument");
  std::vector<BatchLaunch> runs;
  int rc = prepare_batch(b, runs);
  if (rc) return rc;
  for (auto& j : b->jobs) ES_HIP(hipMemsetAsync(j.d_count, 0, 16, b->stream));
  ES_HIP(hipEventRecord(b->ev0, b->stream));
  if ((rc = launch_batch(b, runs, false))) return rc;
  ES_HIP(hipEventRecord(b->ev1, b->stream));
  ES_HIP(hipStreamSynchronize(b->stream));
  float t = 0.f;
  ES_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
  if (ms) *ms = t;
  if (launches) *launches = (uint32_t)runs.size();
  return ES_OK;
}

int es_batch_capture(es_batch* b, uint32_t* launches) {
  if (!b) return fail(ES_ERR_INVALID, "null argument");
  std::vector<BatchLaunch> runs;
  int rc = prepare_batch(b, runs);
  if (rc) return rc;
  b->drop_graph();
  ES_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  rc = launch_batch(b, runs, true);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  ES_HIP(end);
  b->graph = g;
  ES_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
  if (launches) *launches = (uint32_t)runs.size();
  return ES_OK;
}

int es_batch_replay(es_batch* b, double* ms) {
  if (!b) return fail(ES_ERR_INVALID, "null argument");
  if (!b->graph_exec) return fail(ES_ERR_INVALID, "no captured graph (es_batch_capture after the last es_batch_add)");
  ES_HIP(hipEventRecord(b->ev0, b->stream));
  ES_HIP(hipGraphLaunch(b->graph_exec, b->stream));
  ES_HIP(hipEventRecord(b->ev1, b->stream));
  ES_HIP(hipStreamSynchronize(b->stream));
  float t = 0.f;
  ES_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
  if (ms) *ms = t;
  return ES_OK;
}

int es_batch_collect(es_batch* b, uint32_t job, uint64_t* ids, uint64_t cap, uint64_t* n) {
  if (!b || !n || job >= b->jobs.size()) return fail(ES_ERR_INVALID, "bad job");
  const es_batch::Job& j = b->jobs[job];
  unsigned long long c = 0;
  ES_HIP(hipMemcpy(&c, j.d_count, 8, hipMemcpyDeviceToHost));
  *n = c;
  if (c > j.cap) return fail(ES_ERR_OVERFLOW, "survivor buffer overflow");
  if (ids && c) ES_HIP(hipMemcpy(ids, j.d_out, std::min<uint64_t>(c, cap) * 8, hipMemcpyDeviceToHost));
  return ES_OK;
}

}  // extern "C"

// ------------------------------------------------------------ window scan without allocation
//
// es_engine_scan_window_static: the window scan and collect of an engine for callers that
// cannot allocate device memory (HIP graph capture, serving loops).  The kernels' scratch --
// the survivor counter, the slot of the least residue and its floor, one residue record per
// survivor slot, and the exact test's copy of the plan record that points at them -- lives in
// one file-scope device array, `scratch`; survivor IDs go to the caller's device buffer and
// the count and minimum to the caller's device result.  Every GPU holds its own copy of the
// array until the process ends; calls that overlap on one GPU race on it.

namespace {

constexpr uint64_t kStaticMaxIds = 65536;      // survivor slots the scratch records
constexpr size_t kStaticColdOff = 0;           // ColdPlan of the call (the exact test's view)
constexpr size_t kStaticRowOff = 256;          // RowPlan of the call (row-lane engines)
constexpr size_t kStaticCtrOff = 768;          // count, 1 + slot of the minimum, floor
constexpr size_t kStaticRecOff = 1024;         // [kStaticMaxIds][2] residues (lo, hi)
constexpr size_t kScratchBytes = kStaticRecOff + kStaticMaxIds * 16;
static_assert(sizeof(ColdPlan) <= kStaticRowOff - kStaticColdOff, "ColdPlan outgrew its slot");
static_assert(sizeof(RowPlan) <= kStaticCtrOff - kStaticRowOff, "RowPlan outgrew its slot");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[64] = {};         // device address of `scratch`, per device

// Device address of this device's scratch array, looked up once per device (the array's
// name is a host-side stand-in, never a device address).
int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= 64) return fail(ES_ERR_DEVICE, "device index outside 0..63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    ES_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ES_ERR_DEVICE, "hipGetSymbolAddress returned no address for scratch");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return ES_OK;
}

// The call's plan records, zeroed counters and the minimum's floor: written by a kernel (a
// copy or memset node would not be safe to capture).
__global__ void es_static_setup_kernel(ColdPlan cold, RowPlan rows, unsigned char* __restrict__ base,
                                       uint64_t floor) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *reinterpret_cast<ColdPlan*>(base + kStaticColdOff) = cold;
    rows.cold = reinterpret_cast<const ColdPlan*>(base + kStaticColdOff);
    *reinterpret_cast<RowPlan*>(base + kStaticRowOff) = rows;
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(base + kStaticCtrOff);
    ctr[0] = 0;
    ctr[1] = 0;
    ctr[2] = floor;
  }
}

// result = {survivors (may exceed cap), least s lo, least s hi, 1 if there is a minimum}.
__global__ void es_static_finish_kernel(const unsigned char* __restrict__ base,
                                        uint64_t* __restrict__ result) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const unsigned long long* ctr =
        reinterpret_cast<const unsigned long long*>(base + kStaticCtrOff);
    const unsigned long long* rec =
        reinterpret_cast<const unsigned long long*>(base + kStaticRecOff);
    const uint64_t best = ctr[1];
    result[0] = ctr[0];
    result[1] = best ? rec[2 * (best - 1)] : 0;
    result[2] = best ? rec[2 * (best - 1) + 1] : 0;
    result[3] = best ? 1 : 0;
  }
}

}  // namespace

extern "C" int es_engine_scan_window_static(es_engine* e, uint64_t idx_lo, uint64_t idx_hi,
                                            uint64_t* ids, uint64_t cap, uint64_t* result,
                                            void* stream) {
  if (!e || !result || (cap && !ids)) return fail(ES_ERR_INVALID, "null argument");
  if (idx_hi > (uint64_t)e->R || idx_lo > idx_hi) return fail(ES_ERR_INVALID, "window outside [0, R)");
  if (cap > kStaticMaxIds)
    return fail(ES_ERR_LIMIT, "cap exceeds the static scratch: at most 65536 survivor slots");
  if (!e->t_known)
    return fail(ES_ERR_INVALID, "the minimum needs the block index: set it with "
                                "es_engine_set_block (es_engine_set_tconst leaves it unknown)");
  int dev = -1;
  ES_HIP(hipGetDevice(&dev));
  if (dev != e->device) return fail(ES_ERR_INVALID, "the engine's device is not the current one");
  unsigned char* base = nullptr;
  int rc = scratch_address(dev, &base);
  if (rc) return rc;
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : e->stream;
  ColdPlan cold = e->cold;
  cold.out = reinterpret_cast<unsigned long long*>(ids);
  cold.out_n = reinterpret_cast<unsigned long long*>(base + kStaticCtrOff);
  cold.out_s = reinterpret_cast<unsigned long long*>(base + kStaticRecOff);
  cold.out_cap = cap;
  const uint64_t floor = e->t == 0 ? (uint64_t)e->k + 2 : 0;
  hipLaunchKernelGGL(es_static_setup_kernel, dim3(1), dim3(64), 0, st, cold, e->rp, base, floor);
  ES_HIP(hipGetLastError());
  const ColdPlan* cold_dev = reinterpret_cast<const ColdPlan*>(base + kStaticColdOff);
  if (idx_hi > idx_lo) {
    uint64_t o_lo, o_hi;
    outer_range(e, idx_lo, idx_hi, &o_lo, &o_hi);
    const uint64_t n = (o_hi - o_lo + e->launch_outer - 1) / e->launch_outer;
    uint64_t per = (o_hi - o_lo + n - 1) / n;
    per = (per + e->lead - 1) / e->lead * e->lead;
    for (uint64_t a = o_lo; a < o_hi; a += per) {
      rc = launch_kernel(e, st, cold_dev, reinterpret_cast<const RowPlan*>(base + kStaticRowOff),
                         a, std::min(o_hi, a + per), idx_lo, idx_hi, /*no_scratch=*/true);
      if (rc) return rc;
    }
  }
  hipLaunchKernelGGL(es_static_finish_kernel, dim3(1), dim3(64), 0, st,
                     (const unsigned char*)base, result);
  ES_HIP(hipGetLastError());
  return ES_OK;
}
