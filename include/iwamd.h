/*
 * iwamd.h -- extensions of the MI355X ImageWorsener library beyond the
 * reference API in imagew.h.
 *
 * 1. Split-phase processing.  iw_process_image() (src/imagew-main.c:2250-2272)
 *    is prepare -> pixel pipeline -> finish.  iwamd_prepare() runs every host
 *    decision of iw_prepare_processing() (src/imagew-main.c:1899-2248) and
 *    describes the pixel work as a plain-C plan; iwamd_finish() runs what
 *    follows the pixel loops (background label, negate, output optimiser;
 *    src/imagew-main.c:1437-1441, src/imagew-opt.c:1166-1314) on an output
 *    raster supplied by the caller.  iw_process_image() itself is
 *    prepare -> HIP kernels -> finish.  The split lets a scheduler batch the
 *    pixel work of many contexts on the GPU, and lets a test run another
 *    implementation of the pixel pipeline on exactly the same plan.
 *
 * 2. Precision mode.  IWAMD_PRECISION_PLAIN (the default) is the FP64 form
 *    whose output bytes equal the reference's.  IWAMD_PRECISION_FAST resamples
 *    on FP32 MFMA (FP32 operands and sums) and quantises in FP64, for 8- and
 *    16-bit integer outputs; its error against the plain form is bounded and
 *    reported, not zero.
 *
 * 3. Batches and a request loop (iwamd_batch_*, iwamd_server_*).
 */
#ifndef IWAMD_H
#define IWAMD_H

#include <stddef.h>
#include <stdint.h>
#include "imagew.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IWAMD_MAX_CH 4

/* ---- per-dimension resampling settings (src/imagew-internals.h:38-50,
 *      src/imagew-resize.c:412-556) ------------------------------------- */
struct iwamd_resize {
	int family;               /* IW_RESIZETYPE_*, never AUTO after prepare */
	int edge_policy;          /* IW_EDGE_POLICY_* */
	int num_in, num_out;      /* samples along this dimension */
	double param1, param2;    /* as set by iw_set_resize_alg */
	double blur_factor;       /* as set (clamped when weights are built) */
	double out_true_size;
	double translate;
	int use_offset;
	double channel_offset[3]; /* by IW_CHANNELTYPE_RED..BLUE */
};

/* ---- one intermediate channel (src/imagew-internals.h:59-70) ---------- */
struct iwamd_ichan {
	int channeltype;          /* IW_CHANNELTYPE_* */
	int input_channel;        /* first input channel read */
	int cvt_to_grayscale;     /* reads input_channel..+2 */
	int output_channel;       /* -1: resampled but not written */
	int need_unassoc_alpha;   /* premultiply on input, divide on output */
	int early_bkgd;           /* blend with bkgd_color_lin on input */
	double bkgd_color_lin;
	double edge_sample_value[2]; /* virtual-pixel value per dimension */
	struct iw_csdescr in_cs;  /* colour space used to linearise input */
};

/* ---- one output channel (src/imagew-internals.h:72-88) ---------------- */
struct iwamd_ochan {
	int channeltype;
	int ditherfamily, dithersubtype;
	int color_count;          /* 0 = full depth */
	int maxcolorcode_int;
	double maxcolorcode_dbl;
	int use_nc_table;         /* nearest-colour table decides (no dither) */
	int random_seed;          /* PRNG seed when ditherfamily==RANDOM */
	double bkgd1_color_lin, bkgd2_color_lin;
	struct iw_csdescr out_cs; /* colour space used to leave linear light */
};

struct iwamd_plan {
	/* input image (physical layout, src/imagew.h:377-410) */
	const uint8_t *in_pixels;
	size_t in_bpr;
	int in_width, in_height;  /* physical dimensions of img1 */
	int in_bit_depth, in_sampletype, in_imgtype;
	int in_nch_physical;
	unsigned int orient_transform;
	int input_start_x, input_start_y, input_w, input_h;
	int in_alpha_index;       /* logical index of the alpha channel */
	int in_has_alpha;         /* logical input type has alpha */
	double in_maxcolorcode;   /* 2^depth-1 */
	double in_ch_maxcolorcode[IWAMD_MAX_CH];
	int in_ch_slow[IWAMD_MAX_CH]; /* disable_fast_get_sample */
	int grayscale_formula;    /* IW_GSF_WEIGHTED or ORDERBYVALUE */
	double grayscale_weight[3];

	/* intermediate image */
	int inter_nch;
	int inter_alpha_index;    /* -1 when no alpha channel */
	int inter_width, inter_height;
	struct iwamd_ichan ich[IWAMD_MAX_CH];
	struct iwamd_resize rs[2]; /* by IW_DIMENSION_* */
	int intclamp;

	/* output image */
	int out_width, out_height, out_bit_depth, out_sampletype, out_imgtype;
	int out_nch;
	size_t out_bpr;
	struct iwamd_ochan och[IWAMD_MAX_CH];

	/* background (src/imagew-main.c:1734-1823, 1082-1114) */
	int apply_bkgd, bkgd_strategy_late, bkgd_checkerboard, bkgd_check_size;
	int bkgd_check_origin[2];
	int bkgd_has_transparency;
	double bkgd1alpha, bkgd2alpha;

	/* precision form requested for the pixel work */
	int precision;
};

#define IWAMD_PRECISION_PLAIN 0
#define IWAMD_PRECISION_FAST  1

/* Split-phase processing.  iwamd_prepare() consumes the context's single use
 * (a later iw_process_image fails exactly as a second call would).  The plan
 * holds pointers into the context and stays valid until iwamd_finish() or
 * iw_destroy_context().  iwamd_finish() copies out_height*out_bpr bytes from
 * img2 and completes the context so iw_get_output_image() works.  Both return
 * 1 on success, 0 on failure with the context's error flag set. */
IW_EXPORT(int) iwamd_prepare(struct iw_context *ctx, struct iwamd_plan *plan);
IW_EXPORT(int) iwamd_finish(struct iw_context *ctx, const uint8_t *img2);

/* sizeof(struct iwamd_plan), for callers that hold plans opaquely. */
IW_EXPORT(size_t) iwamd_plan_size(void);

/* Precision form for this context (IWAMD_PRECISION_*). */
IW_EXPORT(void) iwamd_set_precision(struct iw_context *ctx, int mode);

/* The quantisation thresholds the GPU path uses in place of libm pow for an
 * output channel with levels 0..P in colour space cstype (gamma for
 * IW_CSTYPE_GAMMA): T[k-1] = least s in [0,1] whose level is >= k (k=1..P),
 * U[k] = least s whose scaled value exceeds k (k=0..P-1), U[P] = +inf.
 * Fills T (P entries) and U (P+1 entries); returns 1. */
IW_EXPORT(int) iwamd_quant_thresholds(int cstype, double gamma, int P, double *T, double *U);

/* The error-diffusion decision data for an output channel (och as in a plan,
 * quantised to P+1 levels): lin[k] the level's value in linear light,
 * M[k] the least s at which a non-exact sample of level k rounds up
 * ((lin[k+1]-s) <= (s-lin[k]) in FP64; 2.0 when never, M[P] = 2.0).  Arrays
 * hold P+1 entries; returns P, or -1 for a floating-point or nearest-colour
 * channel. */
IW_EXPORT(int) iwamd_ed_tables(const struct iwamd_ochan *och, double *lin, double *M);
/* The autotuner's table (no counterpart in the reference).  Band height of
 * the band kernel and tile width of the chosen-precision form change only
 * how the work is cut, not the bytes, so each class of jobs -- form,
 * channels, ceil(log2) of the mean vertical and horizontal taps per output
 * (filter radius x scale factor) and of the output width -- gets the shape
 * that ran fastest on its first tuned job.  Batches tune classes they meet;
 * single jobs (unless the option amd:tune is 1) and the request loop only
 * look them up; IWAMD_TUNE=0 in the environment turns tuning off; the
 * options amd:band_rows (1, 2, 4) and amd:tile (16, 32) fix a job's shape.
 * iwamd_tune_entries() copies up to max entries and returns the count. */
struct iwamd_tune_entry {
	int form;          /* 0: band kernel (shape = output rows per band); 1: chosen-precision tiles (shape = output columns) */
	int channels, taps_v_log2, taps_h_log2, width_log2;
	int shape;         /* the chosen candidate */
	int n;             /* candidates timed */
	int cand[3];
	double ms[3];      /* resampling time per copy of the tuned job, each candidate (copies fill the device) */
};
IW_EXPORT(int) iwamd_tune_entries(struct iwamd_tune_entry *out, int max);

/* The error-diffusion walk's decision for samples s[0..n) in [0, 1]: err[i]
 * and the output level level[i] as the walk computes them, through the level
 * records it keeps in registers for channels of at most 4 levels (tab = 2),
 * the level records and bin index it keeps in LDS (tab = 1) or the bisection
 * it runs over the thresholds for channels with more than 257 levels
 * (tab = 0).  Returns 1, 0 when the channel has more levels than the chosen
 * records hold, -1 for a nearest-colour plan.  No counterpart in the reference: it
 * checks the device's restatement of the decision at
 * src/imagew-main.c:712-748. */
IW_EXPORT(int) iwamd_ed_decide_check(const struct iwamd_ochan *och, int tab, int n, const double *s, double *err,
	int *level);
/* The no-dither quantiser for samples s[0..n) in [0, 1] as the kernels run
 * it (folded boundaries, bin index): code[i].  Returns 1, -1 for a
 * nearest-colour plan.  No counterpart in the reference: it checks the
 * device's restatement of src/imagew-main.c:669-736 with no dither. */
IW_EXPORT(int) iwamd_quant_fold_check(const struct iwamd_ochan *och, int n, const double *s, double *code);

/* An iw_process_image pixel pass that allocates nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  ctx is set up as for iw_process_image, except that the input
 * image's bytes (the layout img1 describes: bpr x rows, 16-bit big-endian)
 * are at d_in in device memory -- img1's host pixels may be NULL -- and the
 * output image is written to d_out in device memory: out_desc's bpr x height
 * bytes (pixels left NULL), as iw_process_image leaves them with every
 * optional optimisation off (iw_set_allow_opt), i.e. the pixel pipeline's
 * image, negated when IW_VAL_NEGATE_TARGET is set, with the colour of fully
 * transparent pixels cleared.  The kernels run on stream (a hipStream_t; NULL:
 * the null stream), and the call returns when they have finished.  Tables,
 * intermediate rows, error-diffusion rows and work lists live in a fixed
 * device array of iwamd_noalloc_scratch_bytes() (64 MiB) per GPU, reserved
 * for the life of the process: a job that needs more fails with an error
 * naming both sizes (C1-sized jobs fit).  Calls on one GPU take turns on
 * that array (a lock held for the call); each call overwrites what the
 * previous one left there.  Never tunes; never calls hipMalloc, hipFree or
 * hipHostMalloc.  Returns 1, or 0 with the context's error set. */
IW_EXPORT(int) iwamd_process_device_noalloc(struct iw_context *ctx, const void *d_in, size_t in_bytes, void *d_out,
	size_t out_bytes, void *stream, struct iw_image *out_desc);
IW_EXPORT(size_t) iwamd_noalloc_scratch_bytes(void);

/* Device-resident jobs.  iwamd_job_create() runs the host decisions on ctx
 * (consuming its single use), builds the device tables and uploads the input
 * image once; iwamd_job_run() launches the pixel kernels on the job's stream
 * over the resident data (repeatable: every run rewrites the output);
 * iwamd_job_sync() waits; iwamd_job_finish() copies the output back and
 * completes ctx so iw_get_output_image() works.  iwamd_job_destroy() frees
 * device memory (it does not destroy ctx).  Returns 0 / NULL on failure with
 * ctx's error flag set. */
struct iwamd_job;
IW_EXPORT(struct iwamd_job*) iwamd_job_create(struct iw_context *ctx);
IW_EXPORT(int) iwamd_job_run(struct iwamd_job *job);
IW_EXPORT(int) iwamd_job_sync(struct iwamd_job *job);
IW_EXPORT(int) iwamd_job_finish(struct iwamd_job *job);
IW_EXPORT(void) iwamd_job_destroy(struct iwamd_job *job);
/* Kernel times of the last run, ms, from HIP events on the job's stream:
 * [0] vertical pass (or the whole fused resample), [1] horizontal pass +
 * quantise (0 when the resample is one fused kernel), [2] error diffusion.
 * Returns the number of entries filled (at most n). */
IW_EXPORT(int) iwamd_job_kernel_ms(struct iwamd_job *job, double *ms, int n);
/* Output samples (pixels) and algorithmic HBM bytes (input + output image)
 * of one run. */
IW_EXPORT(void) iwamd_job_sizes(struct iwamd_job *job, double *out_pixels, double *alg_bytes);
/* Useful floating-point operations of a run's resampling: a multiply and an
 * add per listed tap of both passes, per channel (the chosen form's dense
 * band blocks also multiply the zeros around each list). */
IW_EXPORT(double) iwamd_job_flops(struct iwamd_job *job);

/* Batches.  iwamd_batch_add() runs the host decisions of a context
 * (consuming its single use, like iw_process_image) and stages its tables;
 * iwamd_batch_upload() places every job in one device block (inputs, tables,
 * one output arena, a descriptor table of kernel arguments and the work
 * lists) and captures the launch sequence as a hipGraph: persistent band
 * kernels taking work items from a device counter, then the error-diffusion
 * planes.  iwamd_batch_run() replays the graph (asynchronously; repeatable),
 * iwamd_batch_finish() copies the arena back once and completes every
 * context so iw_get_output_image() works on each.  A job's output bytes do
 * not depend on the other jobs in the batch.  iwamd_batch_add returns the
 * job's index, or -1 with the context's error set; the others return 1 on
 * success, 0 on failure (errors go to the jobs' contexts). */
struct iwamd_batch;
IW_EXPORT(struct iwamd_batch*) iwamd_batch_create(void);
IW_EXPORT(int) iwamd_batch_add(struct iwamd_batch *batch, struct iw_context *ctx);
IW_EXPORT(int) iwamd_batch_upload(struct iwamd_batch *batch);
IW_EXPORT(int) iwamd_batch_run(struct iwamd_batch *batch);
IW_EXPORT(int) iwamd_batch_sync(struct iwamd_batch *batch);
IW_EXPORT(int) iwamd_batch_finish(struct iwamd_batch *batch);
IW_EXPORT(void) iwamd_batch_destroy(struct iwamd_batch *batch);
/* Device time of the last run (HIP events around the graph), ms in ms[0]. */
IW_EXPORT(int) iwamd_batch_kernel_ms(struct iwamd_batch *batch, double *ms, int n);
IW_EXPORT(void) iwamd_batch_sizes(struct iwamd_batch *batch, double *out_pixels, double *alg_bytes, int *n_jobs);
/* As iwamd_job_flops, summed over the batch (after iwamd_batch_upload). */
IW_EXPORT(double) iwamd_batch_flops(struct iwamd_batch *batch);

/* Request loop with continuous batching.  iwamd_server_create() starts a
 * worker thread on the current device with `streams` streams and an HBM
 * budget in bytes (0: 64 GiB).  iwamd_server_submit() runs a context's host
 * decisions on the caller's thread (consuming its single use) and queues it;
 * it returns a request id, or -1 with the context's error set.  The worker
 * admits queued requests while their device bytes fit the budget and
 * launches all requests admitted together as one batch on a free stream;
 * up to `streams` batches overlap on the device.  iwamd_server_wait() blocks
 * until a request is done (1: its context holds the output, 0: failed, with
 * the context's error set); iwamd_server_drain() waits for every request.
 * iwamd_server_request_ms() gives a finished request's submit, start and
 * done times in ms since the server started.  Contexts must outlive the
 * server's use of them (until their request is done).  HIP maps streams onto
 * GPU_MAX_HW_QUEUES hardware queues per process (4 by default): give the
 * process at least `streams` queues (set the variable before the HIP runtime
 * starts), or the streams share queues and a batch's long error-diffusion
 * kernels hold back the batches queued behind them on the same queue. */
struct iwamd_server;
IW_EXPORT(struct iwamd_server*) iwamd_server_create(size_t hbm_budget_bytes, int streams);
IW_EXPORT(int) iwamd_server_submit(struct iwamd_server *server, struct iw_context *ctx);
IW_EXPORT(int) iwamd_server_wait(struct iwamd_server *server, int id);
IW_EXPORT(int) iwamd_server_drain(struct iwamd_server *server);
IW_EXPORT(int) iwamd_server_request_ms(struct iwamd_server *server, int id, double *submit, double *start,
	double *done);
IW_EXPORT(void) iwamd_server_destroy(struct iwamd_server *server);
/* iwamd_server_wait() and iwamd_server_drain() wait at most `ms`
 * milliseconds (default 300000; <= 0: no limit).  On a time-out they return 0
 * and iwamd_server_error() names what is still pending (queued requests,
 * running flights, bytes held); the message is also written to stderr. */
IW_EXPORT(void) iwamd_server_set_timeout(struct iwamd_server *server, double ms);
IW_EXPORT(const char*) iwamd_server_error(struct iwamd_server *server);
/* Device bytes held by running flights now and at most so far, and the
 * number of flights launched. */
IW_EXPORT(void) iwamd_server_stats(struct iwamd_server *server, size_t *bytes_held, size_t *max_bytes_held,
	int *flights);

/* Library/device information.  Returns the number of visible HIP devices
 * (0 when none), and fills name (may be NULL). */
IW_EXPORT(int) iwamd_device_count(void);
/* HBM counter calibration: copies `bytes` (rounded down to 16) between two
 * fresh device buffers with one plain copy kernel (copy_probe_kernel) and
 * reports its time in *ms.  Returns 1 on success.  Profilers read its
 * FETCH_SIZE / WRITE_SIZE against the known byte count. */
IW_EXPORT(int) iwamd_copy_probe(size_t bytes, double *ms);
IW_EXPORT(const char*) iwamd_version(void);

/* Per-stage device timings of the last iw_process_image on this context, in
 * milliseconds, measured with HIP events on the stream the kernels ran on.
 * Fills up to n entries; returns the number of stages. */
IW_EXPORT(int) iwamd_last_timings(struct iw_context *ctx, double *ms, int n);

#ifdef __cplusplus
}
#endif

#endif /* IWAMD_H */
