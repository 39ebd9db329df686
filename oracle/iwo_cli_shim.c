/*
 * iwo_cli_shim.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Linked into the test binary _ref/imagew-oracle (the reference's CLI and
 * codecs built from their own sources), this file defines iw_process_image
 * and so interposes the library's definition for that binary only:
 *
 *   iwamd_prepare()  (the library's host decisions)
 *   iwo_run()        (the oracle's pixel pipeline, on the CPU)
 *   iwamd_finish()   (the library's label / negate / optimiser)
 *
 * The library itself never calls the oracle; this binary exists so the
 * reference's own regression cases can be run through "library decisions +
 * oracle pixels" without a GPU and compared with the reference's outputs.
 */
#include <stdlib.h>

#include "iwo.h"

int iw_process_image(struct iw_context *ctx)
{
	struct iwamd_plan plan;
	uint8_t *img2;
	int ok;
	if (!iwamd_prepare(ctx, &plan)) return 0;
	img2 = malloc(plan.out_bpr * (size_t)plan.out_height + 1);
	if (!img2) return 0;
	ok = iwo_run(&plan, img2);
	if (ok) ok = iwamd_finish(ctx, img2);
	free(img2);
	return ok;
}
