/* es_calib.h -- C ABI of the measurement calibration kernels, part of libes_sieve.so.
 *
 * No counterpart in the reference: the reference prices nothing against hardware peaks.  The
 * sieve's roofline is quoted against the INT32 VALU peak of the device; this kernel measures
 * that peak on the device it runs on, so the quoted peak is confirmed rather than assumed. */
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

#ifdef __cplusplus
}
#endif

#endif /* ES_CALIB_H */
