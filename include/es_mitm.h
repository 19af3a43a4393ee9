/* es_mitm.h -- C ABI of the bounded n-window enumerator (meet in the middle), part of
 * libes_sieve.so.
 *
 * Reference interface replaced: the GPU meet-in-the-middle search of sieve/gpu_mitm.py:32-125
 * (kernel `mitm`, driver mitm_gpu) with the split of harness/mitm_search.py:109-120.  The wheel
 * (whole primes, each at its full modulus q_p) is split into two halves: half 1 is decoded by
 * mixed radix inside the kernel, half 2 is held as a sorted table of its CRT partial sums mod M.
 * For a window [t*M + a, t*M + b) of n (0 <= a < b <= M), each s1 of half 1 meets the circular
 * run of half-2 sums s2 with (s1 + s2) mod M in [a, b), found by binary search; every such n
 * is tested against the remaining primes (filters, full tables mod q_p).  Only n inside the
 * window are formed.  Survivors are the n of the window that pass every filter; with the wheel
 * primes whole, that is every n of the window dominating k for all p <= k. */
#ifndef ES_MITM_H
#define ES_MITM_H

#include <stdint.h>

#include "es_sieve.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct es_mitm es_mitm;

typedef struct es_mitm_desc {
  uint32_t abi_version;        /* ES_ABI_VERSION */
  uint64_t M_lo, M_hi;         /* wheel modulus M < 2^126 */
  uint32_t n1;                 /* rings of half 1 (decoded in the kernel, ring 0 fastest) */
  const uint32_t* radix1;      /* [n1] admissible residues per ring */
  const uint64_t* term1_lo;    /* [sum radix1] CRT terms (a * basis) mod M, ring-major */
  const uint64_t* term1_hi;
  uint32_t n2;                 /* rings of half 2 (the sorted table) */
  const uint32_t* radix2;
  const uint64_t* term2_lo;
  const uint64_t* term2_hi;
  uint32_t n_filters;          /* remaining primes, test order */
  const uint32_t* filt_q;      /* [n_filters] q = p^t_p (< 2^24) */
  const uint32_t* filt_off;    /* [n_filters] word offset of each bitmap in filt_bits */
  const uint32_t* filt_bits;   /* bit r set iff residue r mod q is admissible */
  uint32_t n_filt_words;
  uint64_t out_cap;            /* survivor buffer (n mod M as lo, hi pairs; 0 = 2^20) */
  uint32_t walk;               /* 0: chosen per scan by the expected run length n2*(b-a)/M
                                  (>= 8: wave); 1: one lane per s1; 2: one wave per s1.  In
                                  the sorted sweep: lanes per s1 (4, 8, 16, 32 or 64; 0, 1, 2:
                                  chosen by the expected run length) */
  uint32_t sweep;              /* half 1 as a sorted table too, swept with two pointers (no
                                  per-s1 binary search): 0 when it and its sort buffers fit in
                                  free device memory, 1 never, 2 always */
} es_mitm_desc;

/* Builds and sorts the half-2 table on the device (its size: product of radix2, < 2^32). */
int es_mitm_create(const es_mitm_desc* d, es_mitm** out);
int es_mitm_destroy(es_mitm* m);
/* Sizes: entries of half 1 (decoded, or a sorted table when *swept = 1) and half 2 (table). */
int es_mitm_sizes(es_mitm* m, uint64_t* n1, uint64_t* n2, uint32_t* swept);
/* Survivors of n in [t*M + a, t*M + b), 0 <= a < b <= M (a, b as lo/hi pairs); synchronous.
 * *count may exceed the buffer (then ES_ERR_OVERFLOW). */
int es_mitm_scan(es_mitm* m, uint64_t t, uint64_t a_lo, uint64_t a_hi, uint64_t b_lo,
                 uint64_t b_hi, uint64_t* count);
/* The first `count` survivors as s = n - t*M (lo, hi pairs), unordered. */
int es_mitm_read(es_mitm* m, uint64_t* s_lohi, uint64_t count);
/* Device time of the last scan (ms) and the candidates it tested (n formed in the window). */
int es_mitm_last(es_mitm* m, double* ms, uint64_t* candidates);

#ifdef __cplusplus
}
#endif

#endif /* ES_MITM_H */
