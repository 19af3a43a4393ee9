/* es_calib.h -- C ABI of the measurement calibration kernels, part of libes_sieve.so.
 *
 * No counterpart in the reference: the reference prices nothing against hardware peaks.  These
 * kernels measure, on the device they run on, the peaks the engine's kernels are priced
 * against: the INT32 VALU add rate, the LDS 8-byte read rate and the HBM streaming read rate. */
#ifndef ES_CALIB_H
#define ES_CALIB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runs `blocks` workgroups of 256 threads, each thread `iters` rounds of 8 independent
 * dependent chains of 32 v_add_u32 (256 adds per round), on the current device.  Returns the
 * kernel time (ms, HIP events) and the 32-bit integer adds executed (lanes x adds).  A
 * checksum of the chains is written so the work cannot be removed. */
int es_calib_valu(uint32_t blocks, uint32_t iters, double* ms, uint64_t* ops);

/* Runs `blocks` workgroups of 1024 threads; each thread makes `iters` rounds of 8 reads of an
 * 8-byte row (ds_read_b64) of a 32 KiB LDS table, at its lane's offset plus a wave-uniform
 * step: mode 0 gives a wave 64 consecutive rows (conflict-free: the LDS read peak), mode 1 a
 * fixed pseudo-random row per lane (the sieve's mask-row pattern).  Returns the kernel time
 * (ms, HIP events) and the bytes read from LDS.  The sieve's roofline is priced against the
 * mode-0 rate. */
int es_calib_lds(uint32_t blocks, uint32_t iters, uint32_t mode, double* ms, uint64_t* bytes);

/* Reads `bytes` (a multiple of 16, 16-byte aligned) of device memory at `src` once per launch
 * with a hand-written streaming kernel: `blocks` workgroups of `threads` (256 or 512) threads,
 * each thread loading 16-byte vectors at grid-stride positions, `unroll` (1, 4 or 8) loads in
 * flight, nontemporal when `nontemporal` != 0.  One warm-up launch, then `reps` timed
 * launches (HIP events); returns the best launch time (ms).  The HBM read ceiling the filter
 * pass is compared with. */
int es_calib_hbm_read(const void* src, uint64_t bytes, uint32_t blocks, uint32_t threads,
                      uint32_t unroll, uint32_t nontemporal, uint32_t reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* ES_CALIB_H */
