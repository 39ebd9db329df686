/*
 * iwo.h -- CPU oracle for the ImageWorsener pixel pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg use it, as the checker.
 *
 * iwo_run() executes an iwamd_plan (the decisions of iw_prepare_processing,
 * produced by iwamd_prepare()) the way the reference's iw_process_internal()
 * does (src/imagew-main.c:1368-1461): per channel, column-by-column vertical
 * resample into an FP32 intermediate, then row-by-row horizontal resample,
 * colour-out with libm pow, quantisation, dithering with the sequential PRNG
 * and serpentine error diffusion.  It writes out_height*out_bpr bytes.
 */
#ifndef IWO_H
#define IWO_H

#include <stdint.h>
#include "iwamd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Returns 1 on success, 0 if the plan uses something the oracle lacks. */
int iwo_run(const struct iwamd_plan *plan, uint8_t *img2);

/* Pieces exposed for pinning against closed forms and brute force. */
uint32_t iwo_prng_nth(int seed, uint64_t n);      /* n-th draw (1-based) */
double iwo_filter(int family, double param1, double param2, double x, int num_in, int num_out);
int iwo_weights(const struct iwamd_resize *rs, int channeltype, double edge,
	int *src, double *w, int *start, int cap);    /* returns count or -1 */
double iwo_to_linear(double v, int cstype, double gamma);
double iwo_from_linear(double v, int cstype, double gamma);

#ifdef __cplusplus
}
#endif

#endif
