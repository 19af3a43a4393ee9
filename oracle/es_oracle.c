/* es_oracle.c -- CPU ORACLE for the CRT-wheel sieve.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and
 * only as the checker or as the timed CPU baseline.  The engine never calls it.
 *
 * It restates the reference's CPU Kummer path on the wheel enumeration:
 *   - ref/erdos_ref.py:31-48      digits / dominates (Kummer digit domination)
 *   - ref/erdos_ref.py:69-90      allowed residues per ring and the CRT basis M_i*(M_i^-1 mod q)
 *   - sieve/gpu_sieve_v9.py:127-138  partial_residues (T low digits), ascending order
 *   - sieve/gpu_sieve_v9.py:385-390  idx_to_n: mixed radix, ring 0 fastest, n = t*M + (sum mod M)
 *   - tests/test_enumeration.py:37-43 the per-position loop: idx_to_n, then dominates for every
 *                                   filter prime of the plan.
 * Every table is derived here from (k, wheel items, filter primes) with plain integer loops; no
 * table built by the engine's host code is reused.  n mod p^t_p replaces n in dominates(): the
 * digit loop stops after the t_p digits of k, so only n mod p^t_p is ever read.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define ESO_MAX_RINGS 64
#define ESO_MAX_FILT 512

typedef struct {
  uint32_t k, w, nf;
  uint32_t ring_q[ESO_MAX_RINGS], ring_r[ESO_MAX_RINGS];
  u128* terms[ESO_MAX_RINGS];
  u128 M;
  uint32_t fp[ESO_MAX_FILT], fq[ESO_MAX_FILT], ftc[ESO_MAX_FILT];
} eso_plan;

static uint64_t ipow(uint64_t p, uint32_t e) {
  uint64_t r = 1;
  while (e--) r *= p;
  return r;
}

static uint32_t ndigits(uint32_t k, uint32_t p) {
  uint32_t n = 0;
  while (k) { k /= p; ++n; }
  return n;
}

/* dominates(n, k, p) with n given as n mod p^t_p  (ref/erdos_ref.py:40-48) */
static int dominates_r(uint64_t r, uint32_t k, uint32_t p) {
  while (k) {
    if (k % p > r % p) return 0;
    k /= p;
    r /= p;
  }
  return 1;
}

/* T low base-p digits of r dominate those of k */
static int low_digits_ok(uint64_t r, uint32_t k, uint32_t p, uint32_t T) {
  for (uint32_t i = 0; i < T; ++i) {
    if (k % p > r % p) return 0;
    k /= p;
    r /= p;
  }
  return 1;
}

static int64_t inv_mod(int64_t a, int64_t m) {  /* extended Euclid, gcd(a, m) = 1 */
  int64_t g = m, x = 0, x1 = 1, a1 = a % m;
  if (a1 < 0) a1 += m;
  while (a1) {
    int64_t q = g / a1, t;
    t = g - q * a1; g = a1; a1 = t;
    t = x - q * x1; x = x1; x1 = t;
  }
  if (g != 1) return -1;
  x %= m;
  if (x < 0) x += m;
  return x;
}

static void eso_free(eso_plan* P) {
  for (uint32_t j = 0; j < P->w; ++j) free(P->terms[j]);
  free(P);
}

/* Build the plan: rings in the given order (ring 0 fastest in idx). */
static eso_plan* eso_build(uint32_t k, uint32_t w, const uint32_t* rp, const uint32_t* rT,
                           uint32_t nf, const uint32_t* fp, uint64_t t) {
  if (w > ESO_MAX_RINGS || nf > ESO_MAX_FILT) return NULL;
  eso_plan* P = (eso_plan*)calloc(1, sizeof(eso_plan));
  P->k = k; P->w = w; P->nf = nf;
  P->M = 1;
  for (uint32_t j = 0; j < w; ++j) {
    P->ring_q[j] = (uint32_t)ipow(rp[j], rT[j]);
    P->M *= P->ring_q[j];
  }
  for (uint32_t j = 0; j < w; ++j) {
    const uint32_t p = rp[j], q = P->ring_q[j];
    const u128 Mi = P->M / q;
    const int64_t inv = inv_mod((int64_t)(Mi % q), q);
    if (inv < 0) { P->w = j; eso_free(P); return NULL; }
    uint32_t cnt = 0;
    for (uint32_t r = 0; r < q; ++r) cnt += (uint32_t)low_digits_ok(r, k, p, rT[j]);
    P->ring_r[j] = cnt;
    P->terms[j] = (u128*)malloc(sizeof(u128) * (cnt ? cnt : 1));
    uint32_t i = 0;
    for (uint32_t r = 0; r < q; ++r)
      if (low_digits_ok(r, k, p, rT[j]))  /* ascending residues */
        P->terms[j][i++] = Mi * (u128)(((uint64_t)r * (uint64_t)inv) % q);
  }
  for (uint32_t f = 0; f < nf; ++f) {
    P->fp[f] = fp[f];
    P->fq[f] = (uint32_t)ipow(fp[f], ndigits(k, fp[f]));
    P->ftc[f] = (uint32_t)(((u128)(t % P->fq[f]) * (uint64_t)(P->M % P->fq[f])) % P->fq[f]);
  }
  return P;
}

typedef struct {
  const eso_plan* P;
  uint64_t lo, hi;
  uint64_t* out;
  uint64_t cap, n;
  uint64_t t;
  u128 best;  /* least s over survivors with n = t*M + s > k + 1 (when have) */
  int have;
} eso_job;

static void* eso_worker(void* arg) {
  eso_job* J = (eso_job*)arg;
  const eso_plan* P = J->P;
  for (uint64_t idx = J->lo; idx < J->hi; ++idx) {
    /* idx_to_n: mixed radix, ring 0 fastest */
    uint64_t rem = idx;
    u128 s = 0;
    for (uint32_t j = 0; j < P->w; ++j) {
      const uint64_t d = rem % P->ring_r[j];
      rem /= P->ring_r[j];
      s += P->terms[j][d];
      if (s >= P->M) s -= P->M;
    }
    int ok = 1;
    for (uint32_t f = 0; f < P->nf && ok; ++f) {
      const uint32_t q = P->fq[f];
      const uint64_t r = ((uint64_t)(s % q) + P->ftc[f]) % q;
      ok = dominates_r(r, P->k, P->fp[f]);
    }
    if (ok) {
      if (J->n < J->cap) J->out[J->n] = idx;
      J->n++;
      if ((J->t != 0 || s > (u128)P->k + 1) && (!J->have || s < J->best)) {
        J->best = s;
        J->have = 1;
      }
    }
  }
  return NULL;
}

/* Survivors of positions [idx_lo, idx_hi) of block t, ascending.  Returns the count (which may
 * exceed cap; only cap IDs are written), or -1 on an invalid plan. */
/* As eso_scan, and min_out = {lo, hi, have}: the least s (n = t*M + s) over all survivors with
 * n > k + 1 (ref/erdos_ref.py:121-128: the minimum of a block counts only n > k + 1). */
int64_t eso_scan_min(uint32_t k, uint32_t w, const uint32_t* ring_p, const uint32_t* ring_T,
                     uint32_t nf, const uint32_t* filt_p, uint64_t t, uint64_t idx_lo,
                     uint64_t idx_hi, uint32_t nthreads, uint64_t* out, uint64_t cap,
                     uint64_t min_out[3]);

int64_t eso_scan(uint32_t k, uint32_t w, const uint32_t* ring_p, const uint32_t* ring_T,
                 uint32_t nf, const uint32_t* filt_p, uint64_t t, uint64_t idx_lo,
                 uint64_t idx_hi, uint32_t nthreads, uint64_t* out, uint64_t cap) {
  uint64_t m[3];
  return eso_scan_min(k, w, ring_p, ring_T, nf, filt_p, t, idx_lo, idx_hi, nthreads, out, cap, m);
}

int64_t eso_scan_min(uint32_t k, uint32_t w, const uint32_t* ring_p, const uint32_t* ring_T,
                     uint32_t nf, const uint32_t* filt_p, uint64_t t, uint64_t idx_lo,
                     uint64_t idx_hi, uint32_t nthreads, uint64_t* out, uint64_t cap,
                     uint64_t min_out[3]) {
  eso_plan* P = eso_build(k, w, ring_p, ring_T, nf, filt_p, t);
  if (!P) return -1;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  const uint64_t span = idx_hi > idx_lo ? idx_hi - idx_lo : 0;
  if (span < (uint64_t)nthreads * 1024) nthreads = 1;
  eso_job* jobs = (eso_job*)calloc(nthreads, sizeof(eso_job));
  pthread_t* th = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
  for (uint32_t i = 0; i < nthreads; ++i) {
    jobs[i].P = P;
    jobs[i].lo = idx_lo + span * i / nthreads;
    jobs[i].hi = idx_lo + span * (i + 1) / nthreads;
    jobs[i].cap = cap;
    jobs[i].t = t;
    jobs[i].out = (uint64_t*)malloc(sizeof(uint64_t) * (cap ? cap : 1));
    pthread_create(&th[i], NULL, eso_worker, &jobs[i]);
  }
  uint64_t total = 0;
  u128 best = 0;
  int have = 0;
  for (uint32_t i = 0; i < nthreads; ++i) {
    pthread_join(th[i], NULL);
    if (jobs[i].have && (!have || jobs[i].best < best)) {
      best = jobs[i].best;
      have = 1;
    }
    const uint64_t take = jobs[i].n < cap ? jobs[i].n : cap;
    for (uint64_t j = 0; j < take && total + j < cap; ++j) out[total + j] = jobs[i].out[j];
    total += jobs[i].n;
    free(jobs[i].out);
  }
  free(jobs);
  free(th);
  eso_free(P);
  min_out[0] = (uint64_t)best;
  min_out[1] = (uint64_t)(best >> 64);
  min_out[2] = (uint64_t)have;
  return (int64_t)total;
}

/* Plan geometry for cross-checks: M (lo, hi), R, and the CRT term of (ring j, slot i). */
int eso_geometry(uint32_t k, uint32_t w, const uint32_t* ring_p, const uint32_t* ring_T,
                 uint64_t M_out[2], uint64_t* R_out, uint32_t* radix_out) {
  eso_plan* P = eso_build(k, w, ring_p, ring_T, 0, NULL, 0);
  if (!P) return -1;
  M_out[0] = (uint64_t)P->M;
  M_out[1] = (uint64_t)(P->M >> 64);
  u128 R = 1;
  for (uint32_t j = 0; j < w; ++j) {
    radix_out[j] = P->ring_r[j];
    R *= P->ring_r[j];
  }
  *R_out = (uint64_t)R;
  eso_free(P);
  return 0;
}

/* n = t*M + s of position idx (for inverse-map and minimum checks). */
int eso_idx_to_s(uint32_t k, uint32_t w, const uint32_t* ring_p, const uint32_t* ring_T,
                 uint64_t idx, uint64_t s_out[2]) {
  eso_plan* P = eso_build(k, w, ring_p, ring_T, 0, NULL, 0);
  if (!P) return -1;
  u128 s = 0;
  for (uint32_t j = 0; j < w; ++j) {
    const uint64_t d = idx % P->ring_r[j];
    idx /= P->ring_r[j];
    s += P->terms[j][d];
    if (s >= P->M) s -= P->M;
  }
  s_out[0] = (uint64_t)s;
  s_out[1] = (uint64_t)(s >> 64);
  eso_free(P);
  return 0;
}

/* Kummer test of one n (< 2^128) for one prime: dominates(n, k, p). */
int eso_dominates(uint64_t n_lo, uint64_t n_hi, uint32_t k, uint32_t p) {
  u128 n = ((u128)n_hi << 64) | n_lo;
  while (k) {
    if (k % p > (uint32_t)(n % p)) return 0;
    k /= p;
    n /= p;
  }
  return 1;
}

/* Candidate filter pass restated (sieve/es_word_masks.py:87-88, sieve/gpu_sieve_v8.py:316-322):
 * every candidate ID whose value n dominates k in base p for every listed prime, in input order.
 * Returns the survivor count (only cap IDs are written). */
int64_t eso_filter_values(uint32_t k, uint32_t nf, const uint32_t* primes,
                          const uint64_t* values_lohi, uint64_t n_values, const uint32_t* ids,
                          uint64_t n_ids, uint32_t* out, uint64_t cap) {
  uint64_t qs[ESO_MAX_FILT];
  if (nf > ESO_MAX_FILT) return -1;
  for (uint32_t j = 0; j < nf; ++j) qs[j] = ipow(primes[j], ndigits(k, primes[j]));
  uint64_t n = 0;
  for (uint64_t i = 0; i < n_ids; ++i) {
    if (ids[i] >= n_values) return -1;
    const u128 v = ((u128)values_lohi[2 * (uint64_t)ids[i] + 1] << 64) | values_lohi[2 * (uint64_t)ids[i]];
    int ok = 1;
    for (uint32_t j = 0; j < nf && ok; ++j) ok = dominates_r((uint64_t)(v % qs[j]), k, primes[j]);
    if (ok) {
      if (n < cap) out[n] = ids[i];
      ++n;
    }
  }
  return (int64_t)n;
}

/* Every n in [n_lo, n_hi) that dominates k in base p for every prime p <= k (the Kummer test of
 * ref/erdos_ref.py:40-55 without the n > k+1 condition), ascending: the exact survivor set of an
 * n-space window whatever wheel enumerates it.  pthreads over equal slices.  Returns the count
 * (only cap values are written, lo/hi u64 pairs), or -1 on bad arguments. */
typedef struct {
  uint32_t k, np;
  const uint32_t* primes;
  u128 lo, hi;
  u128* out;
  uint64_t cap, n;
} eso_range_job;

static void* eso_range_worker(void* arg) {
  eso_range_job* J = (eso_range_job*)arg;
  for (u128 n = J->lo; n < J->hi; ++n) {
    int ok = 1;
    for (uint32_t i = 0; i < J->np && ok; ++i) {
      const uint32_t p = J->primes[i];
      uint32_t kk = J->k;
      u128 m = n;
      while (kk) {
        if (kk % p > (uint32_t)(m % p)) { ok = 0; break; }
        kk /= p;
        m /= p;
      }
    }
    if (ok) {
      if (J->n < J->cap) J->out[J->n] = n;
      J->n++;
    }
  }
  return NULL;
}

int64_t eso_good_range(uint32_t k, uint64_t lo_lo, uint64_t lo_hi, uint64_t hi_lo, uint64_t hi_hi,
                       uint32_t nthreads, uint64_t* out_lohi, uint64_t cap) {
  const u128 lo = ((u128)lo_hi << 64) | lo_lo, hi = ((u128)hi_hi << 64) | hi_lo;
  if (hi < lo || k < 2) return -1;
  uint32_t primes[ESO_MAX_FILT], np = 0;
  for (uint32_t p = 2; p <= k && np < ESO_MAX_FILT; ++p) {
    int pr = 1;
    for (uint32_t d = 2; d * d <= p; ++d)
      if (p % d == 0) { pr = 0; break; }
    if (pr) primes[np++] = p;
  }
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  const u128 span = hi - lo;
  if (span < (u128)nthreads * 4096) nthreads = 1;
  eso_range_job* jobs = (eso_range_job*)calloc(nthreads, sizeof(eso_range_job));
  pthread_t* th = (pthread_t*)calloc(nthreads, sizeof(pthread_t));
  for (uint32_t i = 0; i < nthreads; ++i) {
    jobs[i].k = k;
    jobs[i].np = np;
    jobs[i].primes = primes;
    jobs[i].lo = lo + span * i / nthreads;
    jobs[i].hi = lo + span * (i + 1) / nthreads;
    jobs[i].cap = cap;
    jobs[i].out = (u128*)malloc(sizeof(u128) * (cap ? cap : 1));
    pthread_create(&th[i], NULL, eso_range_worker, &jobs[i]);
  }
  uint64_t total = 0;
  for (uint32_t i = 0; i < nthreads; ++i) {
    pthread_join(th[i], NULL);
    for (uint64_t j = 0; j < jobs[i].n && j < cap && total + j < cap; ++j) {
      out_lohi[2 * (total + j)] = (uint64_t)jobs[i].out[j];
      out_lohi[2 * (total + j) + 1] = (uint64_t)(jobs[i].out[j] >> 64);
    }
    total += jobs[i].n;
    free(jobs[i].out);
  }
  free(jobs);
  free(th);
  return (int64_t)total;
}
