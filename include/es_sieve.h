/* es_sieve.h -- C ABI of the MI355X (gfx950) CRT-wheel sieve engine for the Erdos-Selfridge
 * function g(k).
 *
 * One engine = one wheel plan on one GPU, one HIP stream.  The caller describes the plan with
 * plain arrays (ring residues and CRT terms, filter admissibility bitmaps, word-mask
 * prefilters); the library builds every derived device table itself.  A scan enumerates the
 * residue positions idx of block t (n in [tM, (t+1)M)), keeps every idx whose n passes all
 * filters, and appends it to the engine's survivor buffer.
 *
 * Reference interface replaced: the duck-typed engine object of sieve/gpu_sieve_v9.py:425-568
 * (GpuSieve: constructor, tconst.set, out_n.fill(0), _launch(ostart, ocount), out_n.get(),
 * out[:n].get()) and its word-mask variant sieve/es_word_masks.py:95-162.  Survivor IDs are
 * block-relative mixed-radix indices in the ORIGINAL ring order (ring 0 fastest), exactly as the
 * reference emits them (sieve/gpu_sieve_v9.py:325, sieve/es_word_masks.py:43-44,88).
 *
 * Rules: every call returns 0 on success or a negative es_status; no exception crosses the
 * ABI; es_last_error() gives the message of the last failure on the calling thread.  The
 * caller owns every array it passes (they are copied during es_engine_create) and every
 * output buffer.  An engine must be used from one host thread at a time.  Scans are
 * asynchronous; es_engine_count/read/collect/sync synchronise the engine's stream.
 */
#ifndef ES_SIEVE_H
#define ES_SIEVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ES_ABI_VERSION 3u

typedef enum es_status {
  ES_OK = 0,
  ES_ERR_INVALID = -1,    /* bad argument or inconsistent plan (reference: AssertionError) */
  ES_ERR_DEVICE = -2,     /* HIP runtime failure, or no usable GPU */
  ES_ERR_OVERFLOW = -3,   /* survivor buffer overflow (reference: RuntimeError) */
  ES_ERR_LIMIT = -4       /* plan exceeds a documented limit of this engine */
} es_status;

typedef struct es_engine es_engine;

/* Plan descriptor.  Rings are in the reference's ring order (Plan.wheel after the
 * (-radix, p) sort, sieve/gpu_sieve_v9.py:366); ring 0 varies fastest in idx.  An empty
 * wheel is passed as one ring with modulus 1, radix 1, residue 0, term 0.
 *
 * A ring mod p^T may be passed as digit rings instead: its ascending admissible residues are
 * a mixed radix over the base-p digits (digit 0 fastest, digit j ranging over [a_j, p)), so
 * consecutive groups of digits, lowest first, are rings of their own with the same idx order.
 * Each group is passed with modulus p^(digits in the group) and its terms (rho * basis) mod M,
 * rho = the group's digit value sum_j d_j p^j.  When ring 0 is such a group, inner_mod is the
 * whole ring's p^T: ring-0 terms are then multiples of M / inner_mod, the split the kernel
 * uses.
 *
 * The reference's ring j >= 1 may be passed first instead (lead_radix > 1): the rings are then
 * in the order j, 0, 1, .., j-1, j+1, .. and lead_radix is the product of the reference's
 * radices 0..j-1 (ring_radix[1..j] here, the lead block).  idx keeps the reference's order,
 * idx = d + lead_radix * (i + r0 * v) for the lead block's digits d (mixed radix), ring-0
 * slot i of the kernel and the remaining digits v; windows, es_engine_scan units (the
 * reference's ring-0 radix, ring_radix[1], positions each) and es_engine_collect follow it.  A reference ring 0 of more
 * than 32 residues then gives way to a narrower kernel ring 0 with one-word mask rows. */
typedef struct es_plan_desc {
  uint32_t abi_version;        /* ES_ABI_VERSION */
  uint32_t k;
  uint32_t n_rings;            /* w >= 1 */
  const uint32_t* ring_mod;    /* [w] q_j = p^T (< 2^32) */
  const uint32_t* ring_radix;  /* [w] number of admissible residues of ring j */
  const uint64_t* term_lo;     /* [sum radix] CRT term (a * basis_j) mod M, ring-major, */
  const uint64_t* term_hi;     /*             residues of a ring in ascending order     */
  uint64_t M_lo, M_hi;         /* wheel modulus M < 2^126 */
  uint32_t n_filters;          /* filter primes in kernel order (Plan.rest) */
  const uint32_t* filt_q;      /* [n_filters] q = p^t_p  (< 2^24) */
  const uint32_t* filt_off;    /* [n_filters] word offset of each bitmap in filt_bits */
  const uint32_t* filt_bits;   /* bit r of table j set iff residue r mod q_j is admissible */
  uint32_t n_filt_words;       /* total words in filt_bits */
  uint32_t n_slow;             /* filters evaluated exactly per candidate, in this order */
  const uint32_t* slow_order;  /* [n_slow] indices into the filter list */
  uint32_t n_masks;            /* word-mask prefilters: necessary conditions only */
  const uint32_t* mask_mod;    /* [n_masks] modulus m (a divisor of its filter's q) */
  const uint32_t* mask_src;    /* [n_masks] index of the filter the mask derives from */
  const uint32_t* mask_off;    /* [n_masks] word offset of each bitmap in mask_bits */
  const uint32_t* mask_bits;   /* bit r set iff residue r mod m passes the mask */
  uint32_t n_mask_words;
  uint32_t strip_target;       /* target outer indices per uniform strip (0 = default) */
  uint32_t threads;            /* workgroup size: 64, 128, 256, 512 or 1024; 0 = the
                                   size of 256, 512, 1024 that keeps the most waves
                                   resident per CU (ties: the smaller) */
  uint64_t out_cap;            /* survivor buffer capacity in u64 IDs (0 = 2^20) */
  uint32_t lds_budget;         /* max bytes of LDS tables per workgroup (0 = 64 KiB) */
  uint32_t global_masks;       /* masks read through L1 instead of LDS: 0, 2 or 4 (r0 <= 32) */
  uint32_t launch_log2;        /* positions per kernel launch, log2 (0 = 37) */
  uint32_t inner_mod;          /* modulus of the ring ring 0 is a digit group of (0 = ring_mod[0]) */
  uint32_t lead_radix;         /* 0 or 1: rings in the reference order; > 1: the product of
                                  ring_radix[1..j] of a lead block (see above) */
  uint32_t layout;             /* 0: row-lane layout where it applies (one-word rows, every
                                  mask in LDS, at most 11 masks), else strip-lane; 1: always
                                  strip-lane; 2: row-lane, also for two-word rows (a ring 0 of
                                  33..64 slots), or ES_ERR_LIMIT.  Speed only: both layouts
                                  return the same survivors. */
} es_plan_desc;

typedef struct es_engine_info {
  uint64_t R_lo, R_hi;         /* positions per block */
  uint64_t R_outer;            /* outer indices per block (R / r0) */
  uint64_t strips;             /* strips per block */
  uint32_t r0, q0;             /* ring-0 radix and modulus */
  uint32_t strip_len;          /* outer indices per strip */
  uint32_t n_masks_used;       /* masks staged in LDS (may be < n_masks if LDS is short) */
  uint32_t lds_bytes;          /* LDS per workgroup: tables and per-wave candidate queues */
  uint32_t words;              /* u32 words per ring-0 mask row */
  uint32_t threads;
  uint32_t device;
  uint32_t global_masks;       /* masks read through L1 */
  uint32_t layout;             /* 1: strip-lane (a lane walks a strip), 2: row-lane (a lane
                                  holds a strip row, a wave walks steps of the next rings) */
  uint32_t steps;              /* row-lane: steps per unit (V); strip-lane: 0 */
  uint32_t groups;             /* row-lane: lane groups of 64 strip rows (G); strip-lane: 0 */
} es_engine_info;

/* Last error message of the calling thread ("" if none).  Never NULL. */
const char* es_last_error(void);

/* Number of visible GPUs and free/total memory of the current one, in MiB.
 * Counterpart of free_vram_mib (sieve/gpu_sieve_v9.py:416-422). */
int es_device_query(int* n_devices, uint64_t* free_mib, uint64_t* total_mib, char* name, uint32_t name_len);

/* Build an engine for a plan on the current HIP device (GpuSieve.__init__,
 * sieve/gpu_sieve_v9.py:426-498).  The survivor counter starts at zero and t = 0. */
int es_engine_create(const es_plan_desc* desc, es_engine** out);
int es_engine_destroy(es_engine* e);
int es_engine_info_get(const es_engine* e, es_engine_info* info);

/* Per-block constants tconst[j] = (t*M) mod q_j for every filter j, as the reference drivers
 * compute and upload them (sieve/gpu_sieve_v9.py:512-514, harness/run_k_masks.py:131-134).
 * The block index itself stays unknown to the engine (see es_engine_collect). */
int es_engine_set_tconst(es_engine* e, const uint32_t* tconst, uint32_t n);
/* Same, computed by the library from the block index t. */
int es_engine_set_block(es_engine* e, uint64_t t);

/* Zero the survivor counter (out_n.fill(0)). */
int es_engine_reset(es_engine* e);
/* Scan outer indices [ostart, ostart+ocount); each covers r0 positions (GpuSieve._launch,
 * sieve/gpu_sieve_v9.py:500-506).  Asynchronous. */
int es_engine_scan(es_engine* e, uint64_t ostart, uint64_t ocount);
/* Scan positions [idx_lo, idx_hi) of the current block.  Asynchronous. */
int es_engine_scan_window(es_engine* e, uint64_t idx_lo, uint64_t idx_hi);
/* Scan windows bounds[2i] <= idx < bounds[2i+1] (i < n_windows) of the current block one after
 * another, each with its own survivors and minimum, as es_engine_reset + es_engine_scan_window
 * + es_engine_collect would per window; synchronous.  Window i + 1 is queued before window i's
 * results are read back (two result slots of out_cap survivors each, set up on the first call),
 * so the device runs the windows back to back.  The survivor IDs of all windows go to ids in
 * window order (at most cap in total), counts[i] = survivors of window i, min_s[2i..2i+1] =
 * least residue s (n = tM + s > k+1) of window i and has_min[i] = 1 if it has one (min_s and
 * has_min may be NULL; with either given the block index must be known).  ES_ERR_OVERFLOW when
 * a window exceeds out_cap or the IDs exceed cap; counts stay valid.  The engine's own survivor
 * buffer (es_engine_count / es_engine_collect) is left as it was.  No reference counterpart:
 * the reference reads back after every block (sieve/gpu_sieve_v9.py:545-568); this is for
 * callers that walk many windows of one block. */
int es_engine_scan_windows(es_engine* e, const uint64_t* bounds, uint32_t n_windows,
                           uint64_t* ids, uint64_t cap, uint64_t* counts, uint64_t* min_s,
                           uint8_t* has_min);
/* Synchronise; *n = survivors appended since the last reset (may exceed out_cap). */
int es_engine_count(es_engine* e, uint64_t* n);
/* Synchronise; copy survivor IDs [offset, offset+count) (count <= out_cap - offset). */
int es_engine_read(es_engine* e, uint64_t offset, uint64_t count, uint64_t* ids);
/* Synchronise; copy up to cap survivor IDs into ids, set *n to the count, and set
 * min_s = {lo, hi} to the least residue s (n = tM + s) over survivors with n > k+1
 * (*has_min = 0 if there is none).  Returns ES_ERR_OVERFLOW if the buffer overflowed.
 * The minimum is kept on the device while the kernels run (each survivor records s beside
 * its ID; one compare-and-swap per wave keeps the slot of the least s), so only the IDs and
 * one 16-byte record are copied back.  It needs the block index: after es_engine_set_tconst
 * (block unknown) asking for min_s or has_min fails with ES_ERR_INVALID. */
int es_engine_collect(es_engine* e, uint64_t* ids, uint64_t cap, uint64_t* n,
                      uint64_t min_s[2], int* has_min);
int es_engine_sync(es_engine* e);
/* Device time of the sieve kernels launched since the last call (HIP events on the engine's
 * stream), their launch count, and the positions they covered. */
int es_engine_kernel_time(es_engine* e, double* ms, uint64_t* launches, uint64_t* positions);

/* Window scan without allocation: scans positions [idx_lo, idx_hi) of the engine's current
 * block (es_engine_set_block) like es_engine_scan_window and writes, on `stream` (NULL: the
 * engine's stream), the survivor IDs to the caller's device buffer ids[0, min(count, cap)) and
 * result[4] = {count (may exceed cap), least residue s lo, hi (n = tM + s > k+1), 1 if there
 * is a minimum else 0} to the caller's device memory.  Asynchronous; allocates nothing: its
 * scratch (counters, the minimum's slot, one 16-byte residue record per survivor slot, the
 * exact test's plan records) is a file-scope device array of 1 MiB + 1 KiB that every GPU
 * holds until the process ends.  cap above ES_STATIC_MAX_IDS fails with ES_ERR_LIMIT.  Calls
 * that overlap on one GPU race on the shared scratch.  Where the engine runs the row-lane kernel
 * instance sized for 8 waves per SIMD (its spills use per-lane scratch memory, which the HIP
 * runtime sets up at launch), this scan runs the 6-wave instance, which has none; the rarely
 * used instances that spill anyway (three- or four-word rows, runtime word counts, 7 or more
 * masks in some forms) keep their scratch.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (HIP graph capture, tight serving loops). */
#define ES_STATIC_MAX_IDS 65536u
int es_engine_scan_window_static(es_engine* e, uint64_t idx_lo, uint64_t idx_hi, uint64_t* ids,
                                 uint64_t cap, uint64_t* result, void* stream);

/* ---- Batched launches over many plans and blocks ---------------------------------------
 * A batch holds jobs (engine, block t, window [idx_lo, idx_hi)); es_batch_run sieves all of
 * them with one launch per kernel shape (mask count, words per ring-0 row, workgroup size):
 * the jobs' strips are laid out as consecutive workgroup ranges of one grid, each workgroup
 * finds its job by binary search over the ranges and loads that plan's LDS tables, and the
 * hardware dispatcher balances uneven jobs at workgroup granularity.  Each job has its own block
 * constants and survivor buffer, so its survivors equal those of es_engine_scan_window on the
 * same engine.  No reference counterpart: the reference runs one plan per process
 * (harness/grind.py) and one launch per chunk (sieve/gpu_sieve_v9.py:500-506). */
typedef struct es_batch es_batch;

int es_batch_create(uint64_t out_cap_per_job, es_batch** out);
int es_batch_destroy(es_batch* b);
/* The engine must outlive the batch. */
int es_batch_add(es_batch* b, es_engine* e, uint64_t t, uint64_t idx_lo, uint64_t idx_hi,
                 uint32_t* job_id);
/* Synchronous; *ms = device time of all launches, *launches = launches issued. */
int es_batch_run(es_batch* b, double* ms, uint32_t* launches);
int es_batch_collect(es_batch* b, uint32_t job, uint64_t* ids, uint64_t cap, uint64_t* n);
/* Block sweeps through a captured graph: es_batch_capture records the counter reset (a kernel
 * node) and every launch of the current jobs into a hipGraph; es_batch_set_block moves a job to
 * another block t (rewrites its block constants in device memory, outside the graph);
 * es_batch_replay runs the graph (synchronous, *ms = device time).  Adding a job drops the
 * graph.  No reference counterpart: the reference relaunches per chunk from Python. */
int es_batch_capture(es_batch* b, uint32_t* launches);
int es_batch_set_block(es_batch* b, uint32_t job, uint64_t t);
/* Every job at once: t[i] for job i (n_jobs = number of jobs), one synchronisation. */
int es_batch_set_blocks(es_batch* b, const uint64_t* t, uint32_t n_jobs);
int es_batch_replay(es_batch* b, double* ms);

/* ---- Candidate filter pass over a 128-bit value stream (HBM-bound) ------------------------
 * Reference interface replaced: the late scalar filters the word-mask kernel runs on the
 * candidates its masks leave, emitting original IDs (sieve/es_word_masks.py:87-92 with the
 * per-filter test of sieve/gpu_sieve_v8.py:316-322).  Here the candidates are an explicit
 * list of u32 IDs into a stream of u128 values resident in HBM; every ID whose value passes
 * every filter table (bit n mod q_j set) is emitted, in unspecified order. */
typedef struct es_fpass es_fpass;

/* Filters in test order (most selective first); tables are u64 words, bit r = residue r;
 * moduli q in [2, 2^18) (ES_ERR_LIMIT otherwise: the residue arithmetic is exact below). */
int es_fpass_create(uint32_t n_filters, const uint32_t* filt_q, const uint32_t* filt_off,
                    const uint64_t* words, uint32_t n_words, uint32_t lds_budget,
                    uint64_t out_cap, es_fpass** out);
int es_fpass_destroy(es_fpass* f);
/* Copy a host stream (values as lo, hi u64 pairs) and candidate IDs to the device. */
int es_fpass_upload(es_fpass* f, const uint64_t* values_lohi, uint64_t n_values,
                    const uint32_t* ids, uint64_t n_ids);
/* Fill a stream of n values on the device from a seed: each value is
 * (CRT combination of one random admissible residue per conditioned modulus) + J * prod(mods)
 * with J uniform in [0, j_span); candidate IDs are 0..n-1 permuted inside tiles of 64. */
int es_fpass_generate(es_fpass* f, uint64_t n, uint64_t seed, uint32_t n_mods,
                      const uint32_t* mods, const uint32_t* res_count, const uint32_t* residues,
                      uint64_t j_span);
/* Run the pass (synchronous); *n_out = survivors (may exceed out_cap: ES_ERR_OVERFLOW). */
int es_fpass_run(es_fpass* f, uint64_t* n_out);
/* Launch shape of es_fpass_run: candidates per lane per round (2, 4 or 8) and the number of
 * leading filters every candidate is tested against before the survivors of a wave are packed
 * for the remaining filters (default 2 and 3).  Speed only: the survivor set is the same. */
int es_fpass_set_shape(es_fpass* f, uint32_t unroll, uint32_t n_first);
int es_fpass_read(es_fpass* f, uint32_t* ids, uint64_t count);
int es_fpass_read_stream(es_fpass* f, uint64_t offset, uint64_t count, uint64_t* values_lohi,
                         uint32_t* ids);
/* Device time of the last es_fpass_run (HIP events on the pass's stream). */
int es_fpass_last_ms(es_fpass* f, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* ES_SIEVE_H */
