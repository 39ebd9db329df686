/*
 * iwo.c -- CPU oracle (restatement) of the ImageWorsener pixel pipeline.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library, never part of it.
 *
 * The plan it executes comes from iwamd_prepare(); from there on this file
 * follows the reference's own loops and formulas, each function citing the
 * reference lines it restates.  Unlike the GPU library it evaluates libm pow
 * per sample (as the reference does), walks the PRNG sequentially and runs
 * error diffusion with the reference's scatter, so it checks the GPU path's
 * thresholds, jump-ahead and gather reformulations rather than repeating
 * them.  Built with -O3 -ffp-contract=off on x86-64 (no FMA).
 */
#include "iwo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.14159265358979323846

/* ---- colour transfer (src/imagew-main.c:28-96, 427-446) ---------------- */

double iwo_to_linear(double v, int cstype, double gamma)
{
	switch (cstype) {
	case IW_CSTYPE_LINEAR: return v;
	case IW_CSTYPE_GAMMA: return pow(v, gamma);
	case IW_CSTYPE_REC709:
		return (v < 4.5 * 0.020) ? v / 4.5 : pow((v + 0.099) / 1.099, 1.0 / 0.45);
	}
	return (v <= 0.04045) ? v / 12.92 : pow((v + 0.055) / (1.055), 2.4);
}

double iwo_from_linear(double v, int cstype, double gamma)
{
	if (v > 0.999999999) return 1.0;
	switch (cstype) {
	case IW_CSTYPE_LINEAR: return v;
	case IW_CSTYPE_GAMMA: return pow(v, 1.0 / gamma);
	case IW_CSTYPE_REC709: return (v < 0.020) ? 4.5 * v : 1.099 * pow(v, 0.45) - 0.099;
	}
	return (v <= 0.0031308) ? 12.92 * v : 1.055 * pow(v, 1.0 / 2.4) - 0.055;
}

/* ---- filters (src/imagew-resize.c:62-185) ------------------------------ */

typedef struct {
	int family;
	double radius, b, c, mix;
} filt;

static double sinc1(double x)
{
	return (x <= 0.000000005) ? 1.0 : sin(PI_D * x) / (PI_D * x);
}

static double filt_eval(const filt *f, double x)
{
	double v;
	switch (f->family) {
	case IW_RESIZETYPE_LANCZOS:
		return (x < f->radius) ? sinc1(x) * sinc1(x / f->radius) : 0.0;
	case IW_RESIZETYPE_HANNING:
		return (x < f->radius) ? sinc1(x) * (0.5 * cos(PI_D * x / f->radius) + 0.5) : 0.0;
	case IW_RESIZETYPE_BLACKMAN:
		return (x < f->radius) ? sinc1(x) * (0.5 * cos(PI_D * x / f->radius) +
			0.08 * cos(2.0 * PI_D * x / f->radius) + 0.42) : 0.0;
	case IW_RESIZETYPE_SINC:
		return (x < f->radius) ? sinc1(x) : 0.0;
	case IW_RESIZETYPE_GAUSSIAN:
		if (x >= 2.0) return 0.0;
		v = exp(-2.0 * x * x) * 0.79788456080286535587989;
		return (x <= 1.999) ? v : 1000.0 * (2.0 - x) * v;
	case IW_RESIZETYPE_TRIANGLE:
		return (x < 1.0) ? 1.0 - x : 0.0;
	case IW_RESIZETYPE_QUADRATIC:
		if (x < 0.5) return 0.75 - x * x;
		return (x < 1.5) ? 0.50 * (x - 1.5) * (x - 1.5) : 0.0;
	case IW_RESIZETYPE_CUBIC:
		if (x < 1.0)
			return ((12.0 - 9.0 * f->b - 6.0 * f->c) * x * x * x +
				(-18.0 + 12.0 * f->b + 6.0 * f->c) * x * x + (6.0 - 2.0 * f->b)) / 6.0;
		if (x < 2.0)
			return ((-f->b - 6.0 * f->c) * x * x * x + (6.0 * f->b + 30.0 * f->c) * x * x +
				(-12.0 * f->b - 48.0 * f->c) * x + (8.0 * f->b + 24.0 * f->c)) / 6.0;
		return 0.0;
	case IW_RESIZETYPE_HERMITE:
		return (x < 1.0) ? 2.0 * x * x * x - 3.0 * x * x + 1.0 : 0.0;
	case IW_RESIZETYPE_BOX:
		return (x <= 0.5) ? 1.0 : 0.0;
	case IW_RESIZETYPE_BOXAVG:
		if (x < 0.4999999) return 1.0;
		return (x <= 0.5000001) ? 0.5 : 0.0;
	case IW_RESIZETYPE_MIX:
		if (x < (0.5 - f->mix / 2.0)) return 1.0;
		return (x < (0.5 + f->mix / 2.0)) ? 0.5 - (x - 0.5) / f->mix : 0.0;
	}
	return 0.0;
}

/* Filter settings from the family (src/imagew-resize.c:435-522). */
static int filt_setup(filt *f, int family, double p1, double p2, int nin, int nout)
{
	memset(f, 0, sizeof(*f));
	f->family = family;
	switch (family) {
	case IW_RESIZETYPE_MIX:
		f->mix = ((double)nout) / nin;
		if (f->mix > 1.0) f->mix = 1.0 / f->mix;
		f->radius = 0.5 + f->mix;
		return 1;
	case IW_RESIZETYPE_BOX: case IW_RESIZETYPE_BOXAVG: f->radius = 0.501; return 1;
	case IW_RESIZETYPE_TRIANGLE: case IW_RESIZETYPE_HERMITE: f->radius = 1.0; return 1;
	case IW_RESIZETYPE_QUADRATIC: f->radius = 1.5; return 1;
	case IW_RESIZETYPE_GAUSSIAN: f->radius = 2.0; return 1;
	case IW_RESIZETYPE_CUBIC:
		f->radius = 2.0;
		f->b = p1 < -100.0 ? -100.0 : (p1 > 100.0 ? 100.0 : p1);
		f->c = p2 < -100.0 ? -100.0 : (p2 > 100.0 ? 100.0 : p2);
		return 1;
	case IW_RESIZETYPE_LANCZOS: case IW_RESIZETYPE_HANNING:
	case IW_RESIZETYPE_BLACKMAN: case IW_RESIZETYPE_SINC:
		f->radius = floor(p1 + 0.5);
		if (f->radius < 2.0) f->radius = 2.0;
		if (f->radius > 100.0) f->radius = 100.0;
		return 1;
	}
	return 0;
}

double iwo_filter(int family, double p1, double p2, double x, int nin, int nout)
{
	filt f;
	if (!filt_setup(&f, family, p1, p2, nin, nout)) return 0.0;
	return filt_eval(&f, x);
}

/* ---- resize contexts (src/imagew-resize.c:240-571) ---------------------- */

typedef struct {
	int family, nin, nout, edge_policy;
	double out_true, offset, blur, edge_value;
	filt f;
	int n, cap;
	int *src, *dst;
	double *w;
} rrctx;

static void rr_add(rrctx *r, int s, int d, double v)
{
	if (v == 0.0) return;
	if (r->n >= r->cap) {
		r->cap = r->cap * 2 + 64;
		r->src = realloc(r->src, sizeof(int) * r->cap);
		r->dst = realloc(r->dst, sizeof(int) * r->cap);
		r->w = realloc(r->w, sizeof(double) * r->cap);
	}
	r->src[r->n] = s;
	r->dst[r->n] = d;
	r->w[r->n] = v;
	r->n++;
}

/* src/imagew-resize.c:412-556 and :240-344. */
static int rr_init(rrctx *r, const struct iwamd_resize *rs, int channeltype, double edge)
{
	int boxhack = 0, o, i, k;
	memset(r, 0, sizeof(*r));
	r->family = rs->family;
	r->nin = rs->num_in;
	r->nout = rs->num_out;
	r->out_true = rs->out_true_size;
	r->edge_policy = rs->edge_policy;
	r->edge_value = edge;
	if (rs->family == IW_RESIZETYPE_NEAREST || rs->family == IW_RESIZETYPE_BOX) boxhack = 1;
	if (rs->family != IW_RESIZETYPE_NULL && rs->family != IW_RESIZETYPE_NEAREST &&
		!filt_setup(&r->f, rs->family, rs->param1, rs->param2, rs->num_in, rs->num_out))
		return 0;
	r->blur = rs->blur_factor;
	if (r->blur < 0.0001) r->blur = 0.0001;
	if (r->blur > 10000.0) r->blur = 10000.0;
	r->offset = rs->translate;
	if (boxhack) r->offset -= 0.00000000001;
	if (rs->use_offset && channeltype >= 0 && channeltype <= 2) r->offset += rs->channel_offset[channeltype];
	if (rs->family == IW_RESIZETYPE_NULL || rs->family == IW_RESIZETYPE_NEAREST) return 1;

	{
		double red = (r->out_true < (double)r->nin) ? ((double)r->nin) / r->out_true : 1.0;
		red *= r->blur;
		for (o = 0; o < r->nout; o++) {
			double ctr = (0.5 + (double)o - r->offset) / r->out_true;
			double pin = ctr * (double)r->nin - 0.5;
			int first = (int)ceil(pin - r->f.radius * red - 0.0001);
			int last = (int)floor(pin + r->f.radius * red + 0.0001);
			int start = r->n, cnt = 0;
			double vsum = 0.0;
			for (i = first; i <= last; i++) {
				double pos, v;
				int s;
				if (r->edge_policy == IW_EDGE_POLICY_STANDARD && (i < 0 || i >= r->nin)) continue;
				pos = (((double)i) - pin) / red;
				v = filt_eval(&r->f, pos < 0 ? -pos : pos);
				if (v == 0.0) continue;
				vsum += v;
				cnt++;
				if (i < 0 || i > r->nin - 1) {
					if (r->edge_policy == IW_EDGE_POLICY_TRANSPARENT) s = -1;
					else s = (i < 0) ? 0 : r->nin - 1;
				} else {
					s = i;
				}
				rr_add(r, s, o, v);
			}
			if (cnt > 0) {
				for (k = start; k < r->n; k++) r->w[k] = (vsum != 0.0) ? r->w[k] / vsum : 0.0;
			}
		}
	}
	return 1;
}

static void rr_free(rrctx *r)
{
	free(r->src);
	free(r->dst);
	free(r->w);
}

/* iw_resize_row_std / _nearest / _null (src/imagew-resize.c:346-410). */
static void rr_row(const rrctx *r, const double *in, double *out)
{
	int i;
	if (r->family == IW_RESIZETYPE_NULL) {
		for (i = 0; i < r->nout; i++) out[i] = (i < r->nin) ? in[i] : 0.0;
		return;
	}
	if (r->family == IW_RESIZETYPE_NEAREST) {
		for (i = 0; i < r->nout; i++) {
			double ctr = (0.5 + (double)i - r->offset) / (double)r->nout;
			int p = (int)floor(ctr * (double)r->nin);
			if (p < 0) p = 0;
			else if (p > r->nin - 1) p = r->nin - 1;
			out[i] = in[p];
		}
		return;
	}
	for (i = 0; i < r->nout; i++) out[i] = 0.0;
	for (i = 0; i < r->n; i++) {
		if (r->src[i] >= 0) out[r->dst[i]] += in[r->src[i]] * r->w[i];
		else out[r->dst[i]] += r->edge_value * r->w[i];
	}
}

int iwo_weights(const struct iwamd_resize *rs, int channeltype, double edge, int *src, double *w,
	int *start, int cap)
{
	rrctx r;
	int i, n;
	if (!rr_init(&r, rs, channeltype, edge)) return -1;
	n = r.n;
	if (n <= cap) {
		for (i = 0; i <= rs->num_out; i++) start[i] = 0;
		for (i = 0; i < n; i++) {
			src[i] = r.src[i];
			w[i] = r.w[i];
			start[r.dst[i] + 1]++;
		}
		for (i = 0; i < rs->num_out; i++) start[i + 1] += start[i];
	}
	rr_free(&r);
	return n;
}

/* ---- PRNG (src/imagew-util.c:199-233) ---------------------------------- */

typedef struct { uint32_t mult, carry; } prng;

static void prng_seed(prng *p, int s)
{
	p->mult = ((uint32_t)0x03333333) + s;
	p->carry = ((uint32_t)0x05555555) + s;
}

static uint32_t prng_next(prng *p)
{
	uint64_t x = ((uint64_t)0xfff0bf23) * p->mult + p->carry;
	p->carry = (uint32_t)(x >> 32);
	p->mult = 0xffffffff - (uint32_t)(0xffffffff & x);
	return p->mult;
}

uint32_t iwo_prng_nth(int seed, uint64_t n)
{
	prng p;
	uint32_t r = 0;
	uint64_t i;
	prng_seed(&p, seed);
	for (i = 0; i < n; i++) r = prng_next(&p);
	return r;
}

/* ---- the pipeline ------------------------------------------------------- */

typedef struct {
	const struct iwamd_plan *p;
	uint8_t *out;
	float *inter, *inter_alpha, *final_alpha;
	double *err[3];
	prng rng;
} run;

/* translate_coords (src/imagew-main.c:189-234). */
static void phys(const struct iwamd_plan *p, int x, int y, int *rx, int *ry)
{
	int X = p->input_start_x + x, Y = p->input_start_y + y;
	switch (p->orient_transform) {
	case 0: *rx = X; *ry = Y; break;
	case 1: *rx = p->in_width - 1 - X; *ry = Y; break;
	case 2: *rx = X; *ry = p->in_height - 1 - Y; break;
	case 3: *rx = p->in_width - 1 - X; *ry = p->in_height - 1 - Y; break;
	case 4: *rx = Y; *ry = X; break;
	case 5: *rx = Y; *ry = p->in_width - 1 - X; break;
	case 6: *rx = p->in_height - 1 - Y; *ry = X; break;
	case 7: *rx = p->in_height - 1 - Y; *ry = p->in_width - 1 - X; break;
	default: *rx = 0; *ry = 0;
	}
}

/* get_raw_sample_int (src/imagew-main.c:137-253). */
static unsigned int raw_int(const struct iwamd_plan *p, int x, int y, int ch)
{
	int rx, ry;
	const uint8_t *row;
	phys(p, x, y, &rx, &ry);
	row = p->in_pixels + (size_t)ry * p->in_bpr;
	switch (p->in_bit_depth) {
	case 8: return row[p->in_nch_physical * rx + ch];
	case 16: {
		const uint8_t *q = row + (size_t)(p->in_nch_physical * rx + ch) * 2;
		return ((unsigned int)q[0] << 8) | q[1];
	}
	case 4: return (rx & 1) ? (row[rx / 2] & 0x0f) : (row[rx / 2] >> 4);
	case 2: return (row[rx / 4] >> ((3 - rx % 4) * 2)) & 0x03;
	case 1: return (row[rx / 8] & (1 << (7 - rx % 8))) ? 1 : 0;
	}
	return 0;
}

/* get_raw_sample (src/imagew-main.c:257-276): 0..1, virtual alpha = 1. */
static double raw_sample(const struct iwamd_plan *p, int x, int y, int ch)
{
	if (ch >= p->in_nch_physical) return 1.0;
	if (p->in_sampletype == IW_SAMPLETYPE_FLOATINGPOINT) {
		int rx, ry;
		float f;
		phys(p, x, y, &rx, &ry);
		if (p->in_bit_depth != 32) return 0.0;
		memcpy(&f, p->in_pixels + (size_t)ry * p->in_bpr + (size_t)(p->in_nch_physical * rx + ch) * 4, 4);
		return (double)f;
	}
	return ((double)raw_int(p, x, y, ch)) / p->in_ch_maxcolorcode[ch];
}

static double gray(const struct iwamd_plan *p, double r, double g, double b)
{
	/* src/imagew-main.c:278-305 */
	const double *w = p->grayscale_weight;
	double v0, v1, v2;
	if (p->grayscale_formula == IW_GSF_WEIGHTED) return w[0] * r + w[1] * g + w[2] * b;
	if (p->grayscale_formula != IW_GSF_ORDERBYVALUE) return 0.0;
	if (g <= r) { v0 = r; v1 = g; } else { v0 = g; v1 = r; }
	if (b <= v1) v2 = b;
	else {
		v2 = v1;
		if (b <= v0) v1 = b;
		else { v1 = v0; v0 = b; }
	}
	return w[0] * v0 + w[1] * v1 + w[2] * v2;
}

/* get_sample_cvt_to_linear (src/imagew-main.c:307-383). */
static double sample_cvt_to_linear(const struct iwamd_plan *p, const struct iwamd_ichan *ic, int x, int y)
{
	int ch = ic->input_channel, cs = ic->in_cs.cstype;
	double g = ic->in_cs.gamma;
	if (p->in_ch_slow[ch]) {
		if (ic->cvt_to_grayscale)
			return gray(p, iwo_to_linear(raw_sample(p, x, y, ch), cs, g),
				iwo_to_linear(raw_sample(p, x, y, ch + 1), cs, g),
				iwo_to_linear(raw_sample(p, x, y, ch + 2), cs, g));
		return iwo_to_linear(raw_sample(p, x, y, ch), cs, g);
	}
	if (ic->cvt_to_grayscale) {
		double r = iwo_to_linear(((double)raw_int(p, x, y, ch)) / p->in_maxcolorcode, cs, g);
		double gg = iwo_to_linear(((double)raw_int(p, x, y, ch + 1)) / p->in_maxcolorcode, cs, g);
		double b = iwo_to_linear(((double)raw_int(p, x, y, ch + 2)) / p->in_maxcolorcode, cs, g);
		return gray(p, r, gg, b);
	}
	return iwo_to_linear(((double)raw_int(p, x, y, ch)) / p->in_maxcolorcode, cs, g);
}

static void clamp_row(double *v, int n)
{
	int i;
	for (i = 0; i < n; i++) {
		if (v[i] < 0.0) v[i] = 0.0;
		else if (v[i] > 1.0) v[i] = 1.0;
	}
}

/* iw_process_cols_to_intermediate (src/imagew-main.c:831-927). */
static int cols_to_intermediate(run *R, int c)
{
	const struct iwamd_plan *p = R->p;
	const struct iwamd_ichan *ic = &p->ich[c];
	int is_alpha = (c == p->inter_alpha_index);
	int i, j;
	double *in = malloc(sizeof(double) * p->input_h), *out = malloc(sizeof(double) * p->inter_height);
	rrctx r;
	if (!rr_init(&r, &p->rs[IW_DIMENSION_V], ic->channeltype, ic->edge_sample_value[IW_DIMENSION_V])) {
		free(in); free(out);
		return 0;
	}
	for (i = 0; i < p->input_w; i++) {
		for (j = 0; j < p->input_h; j++) {
			in[j] = sample_cvt_to_linear(p, ic, i, j);
			if (ic->need_unassoc_alpha) {
				in[j] *= raw_sample(p, i, j, p->in_alpha_index);
			} else if (ic->early_bkgd) {
				double a = raw_sample(p, i, j, p->in_alpha_index);
				in[j] = (a) * (in[j]) + (1.0 - a) * (ic->bkgd_color_lin);
			}
		}
		rr_row(&r, in, out);
		if (p->intclamp) clamp_row(out, p->inter_height);
		for (j = 0; j < p->inter_height; j++) {
			float *dst = is_alpha ? R->inter_alpha : R->inter;
			dst[(size_t)j * p->inter_width + i] = (float)out[j];
		}
	}
	rr_free(&r);
	free(in);
	free(out);
	return 1;
}

/* iw_ordered_dither pattern (src/imagew-main.c:457-484). */
static const float dither_pattern[2][64] = {
	{0.5f/64, 48.5f/64, 12.5f/64, 60.5f/64, 3.5f/64, 51.5f/64, 15.5f/64, 63.5f/64,
	 32.5f/64, 16.5f/64, 44.5f/64, 28.5f/64, 35.5f/64, 19.5f/64, 47.5f/64, 31.5f/64,
	 8.5f/64, 56.5f/64, 4.5f/64, 52.5f/64, 11.5f/64, 59.5f/64, 7.5f/64, 55.5f/64,
	 40.5f/64, 24.5f/64, 36.5f/64, 20.5f/64, 43.5f/64, 27.5f/64, 39.5f/64, 23.5f/64,
	 2.5f/64, 50.5f/64, 14.5f/64, 62.5f/64, 1.5f/64, 49.5f/64, 13.5f/64, 61.5f/64,
	 34.5f/64, 18.5f/64, 46.5f/64, 30.5f/64, 33.5f/64, 17.5f/64, 45.5f/64, 29.5f/64,
	 10.5f/64, 58.5f/64, 6.5f/64, 54.5f/64, 9.5f/64, 57.5f/64, 5.5f/64, 53.5f/64,
	 42.5f/64, 26.5f/64, 38.5f/64, 22.5f/64, 41.5f/64, 25.5f/64, 37.5f/64, 21.5f/64},
	{3.5f/64, 9.5f/64, 17.5f/64, 27.5f/64, 25.5f/64, 15.5f/64, 7.5f/64, 1.5f/64,
	 11.5f/64, 29.5f/64, 37.5f/64, 45.5f/64, 43.5f/64, 35.5f/64, 23.5f/64, 5.5f/64,
	 19.5f/64, 39.5f/64, 51.5f/64, 57.5f/64, 55.5f/64, 49.5f/64, 33.5f/64, 13.5f/64,
	 31.5f/64, 47.5f/64, 59.5f/64, 63.5f/64, 61.5f/64, 53.5f/64, 41.5f/64, 21.5f/64,
	 30.5f/64, 46.5f/64, 58.5f/64, 62.5f/64, 60.5f/64, 52.5f/64, 40.5f/64, 20.5f/64,
	 18.5f/64, 38.5f/64, 50.5f/64, 56.5f/64, 54.5f/64, 48.5f/64, 32.5f/64, 12.5f/64,
	 10.5f/64, 28.5f/64, 36.5f/64, 44.5f/64, 42.5f/64, 34.5f/64, 22.5f/64, 4.5f/64,
	 2.5f/64, 8.5f/64, 16.5f/64, 26.5f/64, 24.5f/64, 14.5f/64, 6.5f/64, 0.5f/64}};

/* Error-diffusion matrices, slot layout of src/imagew-main.c:506-535. */
static const double ed_matrix[8][12] = {
	{7.0/16, 0.0, 0.0, 3.0/16, 5.0/16, 1.0/16, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
	{7.0/48, 5.0/48, 3.0/48, 5.0/48, 7.0/48, 5.0/48, 3.0/48, 1.0/48, 3.0/48, 5.0/48, 3.0/48, 1.0/48},
	{8.0/42, 4.0/42, 2.0/42, 4.0/42, 8.0/42, 4.0/42, 2.0/42, 1.0/42, 2.0/42, 4.0/42, 2.0/42, 1.0/42},
	{8.0/32, 4.0/32, 2.0/32, 4.0/32, 8.0/32, 4.0/32, 2.0/32, 0.0, 0.0, 0.0, 0.0, 0.0},
	{5.0/32, 3.0/32, 2.0/32, 4.0/32, 5.0/32, 4.0/32, 2.0/32, 0.0, 2.0/32, 3.0/32, 2.0/32, 0.0},
	{4.0/16, 3.0/16, 1.0/16, 2.0/16, 3.0/16, 2.0/16, 1.0/16, 0.0, 0.0, 0.0, 0.0, 0.0},
	{2.0/4, 0.0, 0.0, 1.0/4, 1.0/4, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
	{1.0/8, 1.0/8, 0.0, 1.0/8, 1.0/8, 1.0/8, 0.0, 0.0, 0.0, 1.0/8, 0.0, 0.0}};

/* iw_errdiff_dither (src/imagew-main.c:497-567). */
static void errdiff(run *R, int subtype, double err, int x, int y)
{
	const double *m = ed_matrix[subtype <= 7 ? subtype : 0];
	int W = R->p->out_width, f = (y % 2) ? -1 : 1;
	if ((x - f) >= 0 && (x - f) < W) {
		if ((x - 2 * f) >= 0 && (x - 2 * f) < W) {
			R->err[1][x - 2 * f] += err * m[2];
			R->err[2][x - 2 * f] += err * m[7];
		}
		R->err[1][x - f] += err * m[3];
		R->err[2][x - f] += err * m[8];
	}
	R->err[1][x] += err * m[4];
	R->err[2][x] += err * m[9];
	if ((x + f) >= 0 && (x + f) < W) {
		R->err[0][x + f] += err * m[0];
		R->err[1][x + f] += err * m[5];
		R->err[2][x + f] += err * m[10];
		if ((x + 2 * f) >= 0 && (x + 2 * f) < W) {
			R->err[0][x + 2 * f] += err * m[1];
			R->err[1][x + 2 * f] += err * m[6];
			R->err[2][x + 2 * f] += err * m[11];
		}
	}
}

/* cvt_int_sample_to_linear_output (src/imagew-main.c:330-344). */
static double code_to_linear(unsigned int v, const struct iw_csdescr *cs, double M)
{
	if (cs->cstype == IW_CSTYPE_LINEAR) return ((double)v) / M;
	return iwo_to_linear(((double)v) / M, cs->cstype, cs->gamma);
}

/* get_nearest_valid_colors (src/imagew-main.c:570-629). */
static int nearest_valid(double s, const struct iwamd_ochan *oc, double *lf, double *lc,
	double *ff, double *fc)
{
	double cvt = iwo_from_linear(s, oc->out_cs.cstype, oc->out_cs.gamma), e;
	double M = oc->maxcolorcode_dbl;
	unsigned int fi, ci;
	if (oc->color_count == 0) {
		e = cvt * M;
		if (e > M) e = M;
		if (e < 0.0) e = 0.0;
		*ff = floor(e);
		*fc = ceil(e);
	} else {
		double P = (double)(oc->color_count - 1);
		e = cvt * P;
		if (e > P) e = P;
		if (e < 0.0) e = 0.0;
		*ff = floor(0.5000000001 + floor(e) * (M / P));
		*fc = floor(0.5000000001 + ceil(e) * (M / P));
	}
	fi = (unsigned int)(*ff);
	ci = (unsigned int)(*fc);
	if (fi == ci) return 1;
	*lf = code_to_linear(fi, &oc->out_cs, M);
	*lc = code_to_linear(ci, &oc->out_cs, M);
	return 0;
}

/* get_final_sample_using_nc_tbl (src/imagew-main.c:638-666), over the
 * midpoints of iw_make_nearest_color_table (:1197-1236). */
static double nc_lookup(const double *tbl, double s)
{
	unsigned int x = 127, d = 64;
	for (;;) {
		if (x > 254 || tbl[x] > s) x -= d;
		else x += d;
		if (d == 1) return (x > 254 || tbl[x] > s) ? (double)x : (double)(x + 1);
		d /= 2;
	}
}

static void put_raw(run *R, double s, int x, int y, int ch)
{
	const struct iwamd_plan *p = R->p;
	uint8_t *row = R->out + (size_t)y * p->out_bpr;
	if (p->out_bit_depth == 8) {
		row[p->out_nch * x + ch] = (uint8_t)(0.5 + s);
	} else if (p->out_bit_depth == 16) {
		unsigned short v = (unsigned short)(0.5 + s);
		row[(p->out_nch * x + ch) * 2] = (uint8_t)(v >> 8);
		row[(p->out_nch * x + ch) * 2 + 1] = (uint8_t)(v & 0xff);
	}
}

/* put_sample_convert_from_linear (src/imagew-main.c:669-765). */
static void put_sample(run *R, const double *nc, double s, int x, int y, int o)
{
	const struct iwamd_ochan *oc = &R->p->och[o];
	double lf = 0, lc = 0, ff, fc, dfl, dce, full;
	if (s < 0.0) s = 0.0;
	if (s > 1.0) s = 1.0;
	if (oc->use_nc_table) {
		put_raw(R, nc_lookup(nc, s), x, y, o);
		return;
	}
	if (oc->ditherfamily == IW_DITHERFAMILY_ERRDIFF) {
		s += R->err[0][x];
		if (s > 1.0) s = 1.0;
		else if (s < 0.0) s = 0.0;
	}
	if (nearest_valid(s, oc, &lf, &lc, &ff, &fc)) {
		if (oc->ditherfamily == IW_DITHERFAMILY_RANDOM) (void)prng_next(&R->rng);
		put_raw(R, ff, x, y, o);
		return;
	}
	dfl = s - lf;
	dce = lc - s;
	switch (oc->ditherfamily) {
	case IW_DITHERFAMILY_NONE:
		full = (dce <= dfl) ? fc : ff;
		break;
	case IW_DITHERFAMILY_ERRDIFF:
		if (dce <= dfl) {
			errdiff(R, oc->dithersubtype, -dce, x, y);
			full = fc;
		} else {
			errdiff(R, oc->dithersubtype, dfl, x, y);
			full = ff;
		}
		break;
	case IW_DITHERFAMILY_ORDERED:
		full = (dfl / (dfl + dce) >= dither_pattern[oc->dithersubtype ? 1 : 0][(x % 8) + 8 * (y % 8)]) ? fc : ff;
		break;
	case IW_DITHERFAMILY_RANDOM:
		full = (dfl / (dfl + dce) >= ((double)prng_next(&R->rng)) / (double)0xffffffff) ? fc : ff;
		break;
	default:
		full = 0.0;
	}
	put_raw(R, full, x, y, o);
}

/* iw_process_rows_intermediate_to_final (src/imagew-main.c:929-1155). */
static int rows_to_final(run *R, int c)
{
	const struct iwamd_plan *p = R->p;
	const struct iwamd_ichan *ic = &p->ich[c];
	int is_alpha = (c == p->inter_alpha_index);
	int o = ic->output_channel, i, j, z, k;
	int W = p->out_width, using_ed = 0;
	double *in = malloc(sizeof(double) * p->inter_width), *out = malloc(sizeof(double) * W);
	double nc[255];
	const struct iwamd_ochan *oc = (o >= 0) ? &p->och[o] : NULL;
	rrctx r;

	if (oc && oc->use_nc_table) {
		double prev = 0.0, cur;
		for (i = 0; i < 255; i++) {
			cur = iwo_to_linear(((double)(i + 1)) / 255, oc->out_cs.cstype, oc->out_cs.gamma);
			nc[i] = (prev + cur) / 2.0;
			prev = cur;
		}
	}
	if (oc && oc->ditherfamily == IW_DITHERFAMILY_RANDOM) prng_seed(&R->rng, oc->random_seed);
	if (oc && oc->ditherfamily == IW_DITHERFAMILY_ERRDIFF) {
		using_ed = 1;
		for (k = 0; k < 3; k++)
			for (i = 0; i < W; i++) R->err[k][i] = 0.0;
	}
	if (!rr_init(&r, &p->rs[IW_DIMENSION_H], ic->channeltype, ic->edge_sample_value[IW_DIMENSION_H])) {
		free(in); free(out);
		return 0;
	}
	for (j = 0; j < p->inter_height; j++) {
		const float *src = is_alpha ? R->inter_alpha : R->inter;
		for (i = 0; i < p->inter_width; i++) in[i] = src[(size_t)j * p->inter_width + i];
		rr_row(&r, in, out);
		if (p->intclamp) clamp_row(out, W);
		if (is_alpha)
			for (i = 0; i < W; i++) R->final_alpha[(size_t)j * W + i] = (float)out[i];
		if (o < 0) continue;
		for (z = 0; z < W; z++) {
			double t = 0.0, bkalpha = 0.0;
			int alt = 0;
			i = (using_ed && (j % 2)) ? W - 1 - z : z;
			t = out[i];
			if (p->bkgd_checkerboard)
				alt = (((p->bkgd_check_origin[0] + i) / p->bkgd_check_size) % 2) !=
					(((p->bkgd_check_origin[1] + j) / p->bkgd_check_size) % 2);
			if (p->bkgd_has_transparency) bkalpha = alt ? p->bkgd2alpha : p->bkgd1alpha;
			if (ic->need_unassoc_alpha) {
				double a = R->final_alpha[(size_t)j * W + i];
				if (a != 0.0) t /= a;
				if (p->apply_bkgd && p->bkgd_strategy_late) {
					double bk = alt ? oc->bkgd2_color_lin : oc->bkgd1_color_lin;
					if (p->bkgd_has_transparency) t = t * a + bk * bkalpha * (1.0 - a);
					else t = t * a + bk * (1.0 - a);
				}
			} else if (is_alpha && p->bkgd_has_transparency) {
				t = t + bkalpha * (1.0 - t);
			}
			if (p->out_sampletype == IW_SAMPLETYPE_FLOATINGPOINT) {
				float f = (float)t;
				memcpy(R->out + (size_t)j * p->out_bpr + (size_t)(p->out_nch * i + o) * 4, &f, 4);
			} else {
				put_sample(R, nc, t, i, j, o);
			}
		}
		if (using_ed) {
			for (i = 0; i < W; i++) {
				R->err[0][i] = R->err[1][i];
				R->err[1][i] = R->err[2][i];
				R->err[2][i] = 0.0;
			}
		}
	}
	rr_free(&r);
	free(in);
	free(out);
	return 1;
}

int iwo_run(const struct iwamd_plan *p, uint8_t *img2)
{
	run R;
	int c, k, ok = 1;
	size_t inter = (size_t)p->inter_width * p->inter_height;
	memset(&R, 0, sizeof(R));
	R.p = p;
	R.out = img2;
	memset(img2, 0, p->out_bpr * (size_t)p->out_height);
	R.inter = malloc(sizeof(float) * inter);
	if (p->inter_alpha_index >= 0) {
		R.inter_alpha = malloc(sizeof(float) * inter);
		R.final_alpha = malloc(sizeof(float) * (size_t)p->out_width * p->out_height);
	}
	for (k = 0; k < 3; k++) R.err[k] = calloc(p->out_width, sizeof(double));
	/* Alpha first, then the colour channels (src/imagew-main.c:1410-1435). */
	if (p->inter_alpha_index >= 0) {
		ok = cols_to_intermediate(&R, p->inter_alpha_index) && rows_to_final(&R, p->inter_alpha_index);
	}
	for (c = 0; ok && c < p->inter_nch; c++) {
		if (c == p->inter_alpha_index) continue;
		ok = cols_to_intermediate(&R, c) && rows_to_final(&R, c);
	}
	free(R.inter);
	free(R.inter_alpha);
	free(R.final_alpha);
	for (k = 0; k < 3; k++) free(R.err[k]);
	return ok;
}
