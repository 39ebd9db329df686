// Public API of the MI355X ImageWorsener library: context lifecycle, settings,
// errors, memory and the small utility functions codecs use.
//
// Semantics follow the reference's src/imagew-api.c and src/imagew-util.c
// function by function (cited on each group); the state they touch is the
// context in iw_ctx.h.  The pixel work behind iw_process_image() is in
// iw_gpu.hip; the decisions before it in iw_prepare.cpp; the steps after it
// in iw_finish.cpp.
#include "iw_ctx.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <new>
#include <strings.h>

// ---- allocation (src/imagew-util.c:25-137) --------------------------------

static void *default_malloc(void *, unsigned int flags, size_t n)
{
	return (flags & IW_MALLOCFLAG_ZEROMEM) ? calloc(n, 1) : malloc(n);
}

static void default_free(void *, void *mem) { free(mem); }

extern "C" void *iw_malloc_ex(struct iw_context *ctx, unsigned int flags, size_t n)
{
	if (n > ctx->max_malloc) {
		if (!(flags & IW_MALLOCFLAG_NOERRORS)) iw_set_error(ctx, "Out of memory");
		return nullptr;
	}
	void *mem = ctx->mallocfn(ctx->userdata, flags, n);
	if (!mem && !(flags & IW_MALLOCFLAG_NOERRORS)) iw_set_error(ctx, "Out of memory");
	return mem;
}

extern "C" void *iw_malloc(struct iw_context *ctx, size_t n) { return iw_malloc_ex(ctx, 0, n); }

extern "C" void *iw_mallocz(struct iw_context *ctx, size_t n)
{
	return iw_malloc_ex(ctx, IW_MALLOCFLAG_ZEROMEM, n);
}

extern "C" void *iw_malloc_large(struct iw_context *ctx, size_t n1, size_t n2)
{
	if (n2 && n1 > ctx->max_malloc / n2) {
		iw_set_error(ctx, "Image too large to process");
		return nullptr;
	}
	return iw_malloc_ex(ctx, 0, n1 * n2);
}

// The reference allocates its intermediate buffers with iw_malloc_large; this
// library keeps them on the device but enforces the same max_malloc limit so
// the caller sees the same failures with the same messages.
int iwx_alloc_check(iw_context *ctx, size_t n1, size_t n2)
{
	if (n2 && n1 > ctx->max_malloc / n2) {
		iw_set_error(ctx, "Image too large to process");
		return 0;
	}
	if (n1 * n2 > ctx->max_malloc) {
		iw_set_error(ctx, "Out of memory");
		return 0;
	}
	return 1;
}

extern "C" void *iw_realloc_ex(struct iw_context *ctx, unsigned int flags, void *oldmem,
	size_t oldn, size_t n)
{
	if (!oldmem) return iw_malloc_ex(ctx, flags, n);
	if (n > ctx->max_malloc) {
		if (!(flags & IW_MALLOCFLAG_NOERRORS)) iw_set_error(ctx, "Out of memory");
		return nullptr;
	}
	// Always a fresh block; the old one is freed even on failure.
	void *mem = ctx->mallocfn(ctx->userdata, flags, n);
	if (mem) memcpy(mem, oldmem, oldn < n ? oldn : n);
	ctx->freefn(ctx->userdata, oldmem);
	if (!mem && !(flags & IW_MALLOCFLAG_NOERRORS)) iw_set_error(ctx, "Out of memory");
	return mem;
}

extern "C" void *iw_realloc(struct iw_context *ctx, void *m, size_t oldn, size_t n)
{
	return iw_realloc_ex(ctx, 0, m, oldn, n);
}

extern "C" void iw_free(struct iw_context *ctx, void *mem)
{
	if (mem) ctx->freefn(ctx->userdata, mem);
}

extern "C" char *iw_strdup(struct iw_context *ctx, const char *s)
{
	if (!s) return nullptr;
	size_t len = strlen(s);
	char *d = (char *)iw_malloc(ctx, len + 1);
	if (d) memcpy(d, s, len + 1);
	return d;
}

// ---- strings (src/imagew-util.c:158-197) ----------------------------------

extern "C" void iw_strlcpy(char *dst, const char *src, size_t dstlen)
{
	size_t n = strlen(src);
	if (n > dstlen - 1) n = dstlen - 1;
	memcpy(dst, src, n);
	dst[n] = '\0';
}

extern "C" void iw_vsnprintf(char *buf, size_t buflen, const char *fmt, va_list ap)
{
	vsnprintf(buf, buflen, fmt, ap);
	buf[buflen - 1] = '\0';
}

extern "C" void iw_snprintf(char *buf, size_t buflen, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	iw_vsnprintf(buf, buflen, fmt, ap);
	va_end(ap);
}

extern "C" int iw_stricmp(const char *s1, const char *s2) { return strcasecmp(s1, s2); }

extern "C" void iw_zeromem(void *mem, size_t n) { memset(mem, 0, n); }

// ---- translation, warnings, errors (src/imagew-api.c:16-162) --------------

extern "C" void iw_translate(struct iw_context *ctx, unsigned int flags, char *dst,
	size_t dstlen, const char *src)
{
	dst[0] = '\0';
	int done = (ctx && ctx->translate_fn) ? ctx->translate_fn(ctx, flags, dst, dstlen, src) : 0;
	if (!done) iw_strlcpy(dst, src, dstlen);
}

extern "C" void iw_translatev(struct iw_context *ctx, unsigned int flags, char *dst,
	size_t dstlen, const char *fmt, va_list ap)
{
	if (!ctx || !ctx->translate_fn) {
		iw_vsnprintf(dst, dstlen, fmt, ap);
		return;
	}
	char f[IWX_MSG_MAX], g[IWX_MSG_MAX];
	iw_translate(ctx, IW_TRANSLATEFLAG_FORMAT | flags, f, sizeof(f), fmt);
	iw_vsnprintf(g, sizeof(g), f, ap);
	iw_translate(ctx, IW_TRANSLATEFLAG_POSTFORMAT | flags, dst, dstlen, g);
}

extern "C" void iw_translatef(struct iw_context *ctx, unsigned int flags, char *dst,
	size_t dstlen, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	iw_translatev(ctx, flags, dst, dstlen, fmt, ap);
	va_end(ap);
}

extern "C" void iw_warning(struct iw_context *ctx, const char *s)
{
	if (!ctx->warning_fn) return;
	char buf[IWX_MSG_MAX];
	iw_translate(ctx, IW_TRANSLATEFLAG_WARNINGMSG, buf, sizeof(buf), s);
	ctx->warning_fn(ctx, buf);
}

void iwx_warning(iw_context *ctx, const char *s) { iw_warning(ctx, s); }

extern "C" void iw_warningv(struct iw_context *ctx, const char *fmt, va_list ap)
{
	if (!ctx->warning_fn) return;
	char buf[IWX_MSG_MAX];
	iw_translatev(ctx, IW_TRANSLATEFLAG_WARNINGMSG, buf, sizeof(buf), fmt, ap);
	ctx->warning_fn(ctx, buf);
}

extern "C" void iw_warningf(struct iw_context *ctx, const char *fmt, ...)
{
	if (!ctx->warning_fn) return;
	va_list ap;
	va_start(ap, fmt);
	iw_warningv(ctx, fmt, ap);
	va_end(ap);
}

// First error wins (src/imagew-api.c:106-119).
static void record_error(iw_context *ctx, const char *s)
{
	if (ctx->error_flag) return;
	ctx->error_flag = 1;
	if (!ctx->error_msg) {
		ctx->error_msg = (char *)iw_malloc_ex(ctx, IW_MALLOCFLAG_NOERRORS, IWX_MSG_MAX);
		if (!ctx->error_msg) return;
	}
	iw_strlcpy(ctx->error_msg, s, IWX_MSG_MAX);
}

extern "C" void iw_set_error(struct iw_context *ctx, const char *s)
{
	if (ctx->error_flag) return;
	char buf[IWX_MSG_MAX];
	iw_translate(ctx, IW_TRANSLATEFLAG_ERRORMSG, buf, sizeof(buf), s);
	record_error(ctx, buf);
}

extern "C" void iw_set_errorv(struct iw_context *ctx, const char *fmt, va_list ap)
{
	if (ctx->error_flag) return;
	char buf[IWX_MSG_MAX];
	iw_translatev(ctx, IW_TRANSLATEFLAG_ERRORMSG, buf, sizeof(buf), fmt, ap);
	record_error(ctx, buf);
}

extern "C" void iw_set_errorf(struct iw_context *ctx, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	iw_set_errorv(ctx, fmt, ap);
	va_end(ap);
}

extern "C" const char *iw_get_errormsg(struct iw_context *ctx, char *buf, int buflen)
{
	if (ctx->error_msg)
		iw_strlcpy(buf, ctx->error_msg, buflen);
	else
		iw_translate(ctx, IW_TRANSLATEFLAG_ERRORMSG, buf, buflen, "Error message not available");
	return buf;
}

extern "C" int iw_get_errorflag(struct iw_context *ctx) { return ctx->error_flag; }

// ---- image helpers (src/imagew-api.c:165-207) -----------------------------

extern "C" int iw_imgtype_num_channels(int t)
{
	switch (t) {
	case IW_IMGTYPE_RGBA: return 4;
	case IW_IMGTYPE_RGB: return 3;
	case IW_IMGTYPE_GRAYA: return 2;
	}
	return 1;
}

extern "C" size_t iw_calc_bytesperrow(int num_pixels, int bits_per_pixel)
{
	return (size_t)(((num_pixels * bits_per_pixel) + 7) / 8);
}

extern "C" int iw_check_image_dimensions(struct iw_context *ctx, int w, int h)
{
	if (w > ctx->max_width || h > ctx->max_height) {
		iw_set_errorf(ctx, "Image dimensions too large (%d\xc3\x97%d)", w, h);
		return 0;
	}
	if (w < 1 || h < 1) {
		iw_set_errorf(ctx, "Invalid image dimensions (%d\xc3\x97%d)", w, h);
		return 0;
	}
	return 1;
}

extern "C" int iw_is_valid_density(double dx, double dy, int code)
{
	if (dx < 0.0001 || dy < 0.0001) return 0;
	if (dx > 10000000.0 || dy > 10000000.0) return 0;
	if (dx / 10.0 > dy) return 0;
	if (dy / 10.0 > dx) return 0;
	if (code != IW_DENSITY_UNITS_UNKNOWN && code != IW_DENSITY_UNITS_PER_METER) return 0;
	return 1;
}

// ---- context lifecycle (src/imagew-api.c:209-289) -------------------------

extern "C" struct iw_context *iw_create_context(struct iw_init_params *params)
{
	iw_mallocfn_type mfn = default_malloc;
	iw_freefn_type ffn = default_free;
	void *ud = params ? params->userdata : nullptr;
	if (params && params->mallocfn) {
		mfn = params->mallocfn;
		ffn = params->freefn;
	}
	void *mem = mfn(ud, IW_MALLOCFLAG_ZEROMEM, sizeof(iw_context));
	if (!mem) return nullptr;
	iw_context *ctx = new (mem) iw_context();
	ctx->mallocfn = mfn;
	ctx->freefn = ffn;
	if (params) {
		ctx->userdata = params->userdata;
		ctx->caller_api_version = params->api_version;
	}
	iw_make_srgb_csdescr_2(&ctx->img1cs);
	iw_make_srgb_csdescr_2(&ctx->img2cs);
	return ctx;
}

extern "C" void iw_destroy_context(struct iw_context *ctx)
{
	if (!ctx) return;
	for (auto &o : ctx->req.options) {
		iw_free(ctx, o.name);
		iw_free(ctx, o.val);
	}
	if (ctx->img1.pixels) iw_free(ctx, ctx->img1.pixels);
	if (ctx->img2.pixels) iw_free(ctx, ctx->img2.pixels);
	if (ctx->error_msg) iw_free(ctx, ctx->error_msg);
	if (ctx->optctx.tmp_pixels) iw_free(ctx, ctx->optctx.tmp_pixels);
	if (ctx->optctx.palette) iw_free(ctx, ctx->optctx.palette);
	iw_freefn_type ffn = ctx->freefn;
	void *ud = ctx->userdata;
	ctx->~iw_context();
	ffn(ud, ctx);
}

// ---- output queries (src/imagew-api.c:291-349) ----------------------------

extern "C" void iw_get_output_image(struct iw_context *ctx, struct iw_image *img)
{
	memset(img, 0, sizeof(*img));
	const iwx_opt_ctx &o = ctx->optctx;
	img->width = o.width;
	img->height = o.height;
	img->imgtype = o.imgtype;
	img->sampletype = ctx->img2.sampletype;
	img->bit_depth = o.bit_depth;
	img->pixels = (iw_byte *)o.pixelsptr;
	img->bpr = o.bpr;
	img->density_code = ctx->img2.density_code;
	img->density_x = ctx->img2.density_x;
	img->density_y = ctx->img2.density_y;
	img->rendering_intent = ctx->img2.rendering_intent;
	img->has_bkgdlabel = o.has_bkgdlabel;
	double scale = (o.bit_depth == 8) ? 255.0 : 65535.0;
	for (int k = 0; k < 4; k++) img->bkgdlabel.c[k] = ((double)o.bkgdlabel[k]) / scale;
	img->has_colorkey_trns = o.has_colorkey_trns;
	for (int k = 0; k < 3; k++) img->colorkey[k] = o.colorkey[k];
	if (ctx->reduced_output_maxcolor_flag) {
		img->reduced_maxcolors = 1;
		if (IW_IMGTYPE_IS_GRAY(img->imgtype)) {
			img->maxcolorcode[IW_CHANNELTYPE_GRAY] = ctx->img2_ci[0].maxcolorcode_int;
			if (IW_IMGTYPE_HAS_ALPHA(img->imgtype))
				img->maxcolorcode[IW_CHANNELTYPE_ALPHA] = ctx->img2_ci[1].maxcolorcode_int;
		} else {
			img->maxcolorcode[IW_CHANNELTYPE_RED] = ctx->img2_ci[0].maxcolorcode_int;
			img->maxcolorcode[IW_CHANNELTYPE_GREEN] = ctx->img2_ci[1].maxcolorcode_int;
			img->maxcolorcode[IW_CHANNELTYPE_BLUE] = ctx->img2_ci[2].maxcolorcode_int;
			if (IW_IMGTYPE_HAS_ALPHA(img->imgtype))
				img->maxcolorcode[IW_CHANNELTYPE_ALPHA] = ctx->img2_ci[3].maxcolorcode_int;
		}
	}
}

extern "C" void iw_get_output_colorspace(struct iw_context *ctx, struct iw_csdescr *cs)
{
	*cs = ctx->img2cs;
}

extern "C" const struct iw_palette *iw_get_output_palette(struct iw_context *ctx)
{
	return ctx->optctx.palette;
}

// ---- geometry and output settings (src/imagew-api.c:351-447) --------------

extern "C" void iw_set_output_canvas_size(struct iw_context *ctx, int w, int h)
{
	ctx->canvas_width = w;
	ctx->canvas_height = h;
}

extern "C" void iw_set_output_image_size(struct iw_context *ctx, double w, double h)
{
	ctx->req.out_true_width = (w < 0.01) ? 0.01 : w;
	ctx->req.out_true_height = (h < 0.01) ? 0.01 : h;
	ctx->req.out_true_valid = 1;
}

extern "C" void iw_set_input_crop(struct iw_context *ctx, int x, int y, int w, int h)
{
	ctx->input_start_x = x;
	ctx->input_start_y = y;
	ctx->input_w = w;
	ctx->input_h = h;
}

extern "C" void iw_set_output_profile(struct iw_context *ctx, unsigned int n) { ctx->output_profile = n; }

extern "C" void iw_set_output_depth(struct iw_context *ctx, int bps) { ctx->req.output_depth = bps; }

extern "C" void iw_set_output_max_color_code(struct iw_context *ctx, int ct, int n)
{
	if (ct >= 0 && ct < IWX_NUM_CHANNELTYPES) ctx->req.output_maxcolorcode[ct] = n;
}

// IW_CHANNELTYPE_ALL also covers alpha; NONALPHA covers R,G,B and gray
// (src/imagew-api.c:391-433).
static void set_per_channeltype(int *arr, int ct, int v)
{
	if (ct >= 0 && ct < IWX_NUM_CHANNELTYPES) arr[ct] = v;
	if (ct == IW_CHANNELTYPE_ALL) arr[IW_CHANNELTYPE_ALPHA] = v;
	if (ct == IW_CHANNELTYPE_ALL || ct == IW_CHANNELTYPE_NONALPHA) {
		arr[IW_CHANNELTYPE_RED] = v;
		arr[IW_CHANNELTYPE_GREEN] = v;
		arr[IW_CHANNELTYPE_BLUE] = v;
		arr[IW_CHANNELTYPE_GRAY] = v;
	}
}

extern "C" void iw_set_dither_type(struct iw_context *ctx, int ct, int f, int s)
{
	set_per_channeltype(ctx->ditherfamily_by_channeltype, ct, f);
	set_per_channeltype(ctx->dithersubtype_by_channeltype, ct, s);
}

extern "C" void iw_set_color_count(struct iw_context *ctx, int ct, int c)
{
	set_per_channeltype(ctx->req.color_count, ct, c);
}

extern "C" void iw_set_channel_offset(struct iw_context *ctx, int ct, int dimension, double offs)
{
	if (ct < 0 || ct > 2) return;
	if (dimension < 0 || dimension > 1) dimension = 0;
	ctx->resize_settings[dimension].channel_offset[ct] = offs;
}

extern "C" void iw_set_input_max_color_code(struct iw_context *ctx, int input_channel, int c)
{
	if (input_channel >= 0 && input_channel < IWX_CI_COUNT)
		ctx->img1_ci[input_channel].maxcolorcode_int = c;
}

// ---- background colours and labels (src/imagew-api.c:449-608) -------------

extern "C" void iw_set_input_bkgd_label_2(struct iw_context *ctx, const struct iw_color *clr)
{
	ctx->img1_bkgd_label_set = 1;
	ctx->img1_bkgd_label_inputcs = *clr;
}

extern "C" void iw_set_input_bkgd_label(struct iw_context *ctx, double r, double g, double b)
{
	struct iw_color c = {{r, g, b, 1.0}};
	iw_set_input_bkgd_label_2(ctx, &c);
}

extern "C" void iw_set_output_bkgd_label_2(struct iw_context *ctx, const struct iw_color *clr)
{
	ctx->req.output_bkgd_label_valid = 1;
	ctx->req.output_bkgd_label = *clr;
}

extern "C" void iw_set_output_bkgd_label(struct iw_context *ctx, double r, double g, double b)
{
	struct iw_color c = {{r, g, b, 1.0}};
	iw_set_output_bkgd_label_2(ctx, &c);
}

extern "C" void iw_set_apply_bkgd_2(struct iw_context *ctx, const struct iw_color *clr)
{
	ctx->req.bkgd_valid = 1;
	ctx->req.bkgd = *clr;
}

extern "C" void iw_set_apply_bkgd(struct iw_context *ctx, double r, double g, double b)
{
	struct iw_color c = {{r, g, b, 1.0}};
	iw_set_apply_bkgd_2(ctx, &c);
}

extern "C" void iw_set_bkgd_checkerboard_2(struct iw_context *ctx, int checkersize,
	const struct iw_color *clr)
{
	ctx->req.bkgd_checkerboard = 1;
	ctx->bkgd_check_size = checkersize;
	ctx->req.bkgd2 = *clr;
}

extern "C" void iw_set_bkgd_checkerboard(struct iw_context *ctx, int checkersize, double r2,
	double g2, double b2)
{
	struct iw_color c = {{r2, g2, b2, 1.0}};
	iw_set_bkgd_checkerboard_2(ctx, checkersize, &c);
}

extern "C" void iw_set_bkgd_checkerboard_origin(struct iw_context *ctx, int x, int y)
{
	ctx->bkgd_check_origin[IW_DIMENSION_H] = x;
	ctx->bkgd_check_origin[IW_DIMENSION_V] = y;
}

// ---- density (src/imagew-api.c:481-508) -----------------------------------

extern "C" int iw_get_input_density(struct iw_context *ctx, double *px, double *py, int *pcode)
{
	*px = 1.0;
	*py = 1.0;
	*pcode = IW_DENSITY_UNKNOWN;
	if (ctx->img1.density_code == IW_DENSITY_UNKNOWN) return 0;
	if (!iw_is_valid_density(ctx->img1.density_x, ctx->img1.density_y, ctx->img1.density_code))
		return 0;
	*px = ctx->img1.density_x;
	*py = ctx->img1.density_y;
	*pcode = ctx->img1.density_code;
	return 1;
}

extern "C" void iw_set_output_density(struct iw_context *ctx, double x, double y, int code)
{
	ctx->img2.density_code = code;
	ctx->img2.density_x = x;
	ctx->img2.density_y = y;
}

// ---- colour-space descriptors (src/imagew-api.c:510-567) ------------------

// A gamma within 5e-6 of 1 is linear (src/imagew-api.c:511-517).
static void gamma_near_one_is_linear(struct iw_csdescr *cs)
{
	if (cs->cstype == IW_CSTYPE_GAMMA && cs->gamma >= 0.999995 && cs->gamma <= 1.000005)
		cs->cstype = IW_CSTYPE_LINEAR;
}

extern "C" void iw_make_linear_csdescr(struct iw_csdescr *cs)
{
	cs->cstype = IW_CSTYPE_LINEAR;
	cs->gamma = 0.0;
	cs->srgb_intent = 0;
}

extern "C" void iw_make_srgb_csdescr(struct iw_csdescr *cs, int srgb_intent)
{
	cs->cstype = IW_CSTYPE_SRGB;
	cs->gamma = 0.0;
	cs->srgb_intent = srgb_intent;
}

extern "C" void iw_make_srgb_csdescr_2(struct iw_csdescr *cs)
{
	cs->cstype = IW_CSTYPE_SRGB;
	cs->gamma = 0.0;
}

extern "C" void iw_make_rec709_csdescr(struct iw_csdescr *cs)
{
	cs->cstype = IW_CSTYPE_REC709;
	cs->gamma = 0.0;
}

extern "C" void iw_make_gamma_csdescr(struct iw_csdescr *cs, double gamma)
{
	cs->cstype = IW_CSTYPE_GAMMA;
	cs->gamma = gamma < 0.1 ? 0.1 : (gamma > 10.0 ? 10.0 : gamma);
	cs->srgb_intent = 0;
	gamma_near_one_is_linear(cs);
}

extern "C" void iw_set_output_colorspace(struct iw_context *ctx, const struct iw_csdescr *cs)
{
	ctx->req.output_cs = *cs;
	gamma_near_one_is_linear(&ctx->req.output_cs);
	ctx->req.output_cs_valid = 1;
}

extern "C" void iw_set_input_colorspace(struct iw_context *ctx, const struct iw_csdescr *cs)
{
	ctx->img1cs = *cs;
	gamma_near_one_is_linear(&ctx->img1cs);
}

// ---- misc setters (src/imagew-api.c:610-748) ------------------------------

extern "C" void iw_set_max_malloc(struct iw_context *ctx, size_t n) { ctx->max_malloc = n; }

extern "C" void iw_set_random_seed(struct iw_context *ctx, int randomize, int rand_seed)
{
	ctx->randomize = randomize;
	ctx->random_seed = rand_seed;
}

extern "C" void iw_set_userdata(struct iw_context *ctx, void *userdata) { ctx->userdata = userdata; }
extern "C" void *iw_get_userdata(struct iw_context *ctx) { return ctx->userdata; }

extern "C" void iw_set_translate_fn(struct iw_context *ctx, iw_translatefn_type fn)
{
	ctx->translate_fn = fn;
}

extern "C" void iw_set_warning_fn(struct iw_context *ctx, iw_warningfn_type fn) { ctx->warning_fn = fn; }

// Takes ownership of img->pixels (freed in iw_destroy_context).
extern "C" void iw_set_input_image(struct iw_context *ctx, const struct iw_image *img)
{
	ctx->img1 = *img;
}

extern "C" void iw_set_resize_alg(struct iw_context *ctx, int dimension, int family,
	double blur, double param1, double param2)
{
	if (dimension < 0 || dimension > 1) dimension = 0;
	iwx_resize_settings &rs = ctx->resize_settings[dimension];
	rs.family = family;
	rs.blur_factor = blur;
	rs.param1 = param1;
	rs.param2 = param2;
}

// A change of orientation composes with the current transform; a transpose
// swaps the logical dimensions (src/imagew-api.c:660-684).
extern "C" void iw_reorient_image(struct iw_context *ctx, unsigned int x)
{
	static const unsigned int after_transpose[8] = {4, 6, 5, 7, 0, 2, 1, 3};
	x &= 0x07;
	if (x & 0x04) {
		ctx->img1.orient_transform = after_transpose[ctx->img1.orient_transform];
		int t = ctx->img1.width;
		ctx->img1.width = ctx->img1.height;
		ctx->img1.height = t;
		double d = ctx->img1.density_x;
		ctx->img1.density_x = ctx->img1.density_y;
		ctx->img1.density_y = d;
	}
	ctx->img1.orient_transform ^= (x & 0x03);
}

extern "C" int iw_get_sample_size(void) { return (int)sizeof(float); }
extern "C" int iw_get_version_int(void) { return IW_VERSION_INT; }

extern "C" char *iw_get_version_string(struct iw_context *, char *s, int s_len)
{
	int v = iw_get_version_int();
	iw_snprintf(s, s_len, "%d.%d.%d", (v & 0xff0000) >> 16, (v & 0xff00) >> 8, v & 0xff);
	return s;
}

extern "C" char *iw_get_copyright_string(struct iw_context *ctx, char *dst, int dstlen)
{
	iw_translatef(ctx, 0, dst, dstlen, "Copyright \xc2\xa9 %s %s",
		"2011" "\xe2\x80\x93" "2022", "Jason Summers");
	return dst;
}

extern "C" void iw_set_zlib_module(struct iw_context *ctx, struct iw_zlib_module *z) { ctx->zlib_module = z; }
extern "C" struct iw_zlib_module *iw_get_zlib_module(struct iw_context *ctx) { return ctx->zlib_module; }

extern "C" void iw_set_allow_opt(struct iw_context *ctx, int opt, int n)
{
	iw_byte v = n ? 1 : 0;
	switch (opt) {
	case IW_OPT_GRAYSCALE: ctx->opt_grayscale = v; break;
	case IW_OPT_PALETTE: ctx->opt_palette = v; break;
	case IW_OPT_16_TO_8: ctx->opt_16_to_8 = v; break;
	case IW_OPT_STRIP_ALPHA: ctx->opt_strip_alpha = v; break;
	case IW_OPT_BINARY_TRNS: ctx->opt_binary_trns = v; break;
	}
}

// Weights are normalised to sum to one (src/imagew-api.c:735-748).
extern "C" void iw_set_grayscale_weights(struct iw_context *ctx, double r, double g, double b)
{
	double tot = r + g + b;
	if (tot == 0.0) tot = 1.0;
	ctx->grayscale_weight[0] = r / tot;
	ctx->grayscale_weight[1] = g / tot;
	ctx->grayscale_weight[2] = b / tot;
}

extern "C" unsigned int iw_color_get_int_sample(struct iw_color *clr, int channel, unsigned int maxcc)
{
	int n = (int)(0.5 + (clr->c[channel] * (double)maxcc));
	if (n < 0) n = 0;
	else if (n > (int)maxcc) n = (int)maxcc;
	return (unsigned int)n;
}

// ---- generic settings (src/imagew-api.c:760-1042) -------------------------

static char *strdup_number(iw_context *ctx, double n)
{
	char buf[100];
	iw_snprintf(buf, sizeof(buf), "%.20f", n);
	return iw_strdup(ctx, buf);
}

extern "C" void iw_set_value(struct iw_context *ctx, int code, int n)
{
	switch (code) {
	case IW_VAL_API_VERSION: ctx->caller_api_version = n; break;
	case IW_VAL_CVT_TO_GRAYSCALE: ctx->to_grayscale = n; break;
	case IW_VAL_DISABLE_GAMMA: ctx->no_gamma = n; break;
	case IW_VAL_NO_CSLABEL: ctx->req.suppress_output_cslabel = n; break;
	case IW_VAL_INT_CLAMP: ctx->intclamp = n; break;
	case IW_VAL_EDGE_POLICY_X: ctx->resize_settings[IW_DIMENSION_H].edge_policy = n; break;
	case IW_VAL_EDGE_POLICY_Y: ctx->resize_settings[IW_DIMENSION_V].edge_policy = n; break;
	case IW_VAL_PREF_UNITS: ctx->pref_units = n; break;
	case IW_VAL_GRAYSCALE_FORMULA: ctx->grayscale_formula = n; break;
	case IW_VAL_INPUT_NATIVE_GRAYSCALE: ctx->img1.native_grayscale = n; break;
	case IW_VAL_COMPRESSION: ctx->req.compression = n; break;
	case IW_VAL_PAGE_TO_READ: ctx->req.page_to_read = n; break;
	case IW_VAL_INCLUDE_SCREEN: ctx->req.include_screen = n; break;
	case IW_VAL_JPEG_QUALITY: iw_set_option(ctx, "jpeg:quality", strdup_number(ctx, n)); break;
	case IW_VAL_JPEG_SAMP_FACTOR_H: iw_set_option(ctx, "jpeg:sampling-x", strdup_number(ctx, n)); break;
	case IW_VAL_JPEG_SAMP_FACTOR_V: iw_set_option(ctx, "jpeg:sampling-y", strdup_number(ctx, n)); break;
	case IW_VAL_JPEG_ARITH_CODING: iw_set_option(ctx, "jpeg:arith", strdup_number(ctx, n)); break;
	case IW_VAL_DEFLATE_CMPR_LEVEL: iw_set_option(ctx, "deflate:cmprlevel", strdup_number(ctx, n)); break;
	case IW_VAL_OUTPUT_INTERLACED: ctx->req.interlaced = n; break;
	case IW_VAL_USE_BKGD_LABEL: ctx->req.use_bkgd_label_from_file = n; break;
	case IW_VAL_BMP_NO_FILEHEADER: ctx->req.bmp_no_fileheader = n; break;
	case IW_VAL_BMP_VERSION: iw_set_option(ctx, "bmp:version", strdup_number(ctx, n)); break;
	case IW_VAL_MAX_WIDTH: ctx->max_width = n; break;
	case IW_VAL_MAX_HEIGHT: ctx->max_height = n; break;
	case IW_VAL_NO_BKGD_LABEL: ctx->req.suppress_output_bkgd_label = n; break;
	case IW_VAL_INTENT: ctx->req.output_rendering_intent = n; break;
	case IW_VAL_OUTPUT_SAMPLE_TYPE: ctx->req.output_sample_type = n; break;
	case IW_VAL_OUTPUT_COLOR_TYPE:
		if (n == IW_COLORTYPE_RGB) iw_set_option(ctx, "deflate:colortype", "rgb");
		break;
	case IW_VAL_OUTPUT_FORMAT: ctx->req.output_format = n; break;
	case IW_VAL_NEGATE_TARGET: ctx->req.negate_target = n; break;
	}
}

extern "C" int iw_get_value(struct iw_context *ctx, int code)
{
	switch (code) {
	case IW_VAL_API_VERSION: return ctx->caller_api_version;
	case IW_VAL_CVT_TO_GRAYSCALE: return ctx->to_grayscale;
	case IW_VAL_DISABLE_GAMMA: return ctx->no_gamma;
	case IW_VAL_NO_CSLABEL: return ctx->req.suppress_output_cslabel;
	case IW_VAL_INT_CLAMP: return ctx->intclamp;
	case IW_VAL_EDGE_POLICY_X: return ctx->resize_settings[IW_DIMENSION_H].edge_policy;
	case IW_VAL_EDGE_POLICY_Y: return ctx->resize_settings[IW_DIMENSION_V].edge_policy;
	case IW_VAL_PREF_UNITS: return ctx->pref_units;
	case IW_VAL_GRAYSCALE_FORMULA: return ctx->grayscale_formula;
	case IW_VAL_INPUT_NATIVE_GRAYSCALE: return ctx->img1.native_grayscale;
	case IW_VAL_INPUT_WIDTH: return ctx->img1.width < 1 ? 1 : ctx->img1.width;
	case IW_VAL_INPUT_HEIGHT: return ctx->img1.height < 1 ? 1 : ctx->img1.height;
	case IW_VAL_INPUT_IMAGE_TYPE: return ctx->img1.imgtype;
	case IW_VAL_INPUT_DEPTH: return ctx->img1.bit_depth;
	case IW_VAL_COMPRESSION: return ctx->req.compression;
	case IW_VAL_PAGE_TO_READ: return ctx->req.page_to_read;
	case IW_VAL_INCLUDE_SCREEN: return ctx->req.include_screen;
	case IW_VAL_OUTPUT_PALETTE_GRAYSCALE: return ctx->optctx.palette_is_grayscale;
	case IW_VAL_OUTPUT_INTERLACED: return ctx->req.interlaced;
	case IW_VAL_USE_BKGD_LABEL: return ctx->req.use_bkgd_label_from_file;
	case IW_VAL_BMP_NO_FILEHEADER: return ctx->req.bmp_no_fileheader;
	case IW_VAL_MAX_WIDTH: return ctx->max_width;
	case IW_VAL_MAX_HEIGHT: return ctx->max_height;
	case IW_VAL_PRECISION: return 32;
	case IW_VAL_NO_BKGD_LABEL: return ctx->req.suppress_output_bkgd_label;
	case IW_VAL_INTENT: return ctx->req.output_rendering_intent;
	case IW_VAL_OUTPUT_SAMPLE_TYPE: return ctx->req.output_sample_type;
	case IW_VAL_OUTPUT_FORMAT: return ctx->req.output_format;
	case IW_VAL_NEGATE_TARGET: return ctx->req.negate_target;
	}
	return 0;
}

extern "C" void iw_set_value_dbl(struct iw_context *ctx, int code, double n)
{
	switch (code) {
	case IW_VAL_WEBP_QUALITY: iw_set_option(ctx, "webp:quality", strdup_number(ctx, n)); break;
	case IW_VAL_TRANSLATE_X: ctx->resize_settings[IW_DIMENSION_H].translate = n; break;
	case IW_VAL_TRANSLATE_Y: ctx->resize_settings[IW_DIMENSION_V].translate = n; break;
	}
}

extern "C" double iw_get_value_dbl(struct iw_context *ctx, int code)
{
	switch (code) {
	case IW_VAL_TRANSLATE_X: return ctx->resize_settings[IW_DIMENSION_H].translate;
	case IW_VAL_TRANSLATE_Y: return ctx->resize_settings[IW_DIMENSION_V].translate;
	}
	return 0.0;
}

// At most 32 options; an empty value means "1" (src/imagew-api.c:996-1042).
extern "C" void iw_set_option(struct iw_context *ctx, const char *name, const char *val)
{
	if (!val || !val[0]) val = "1";
	for (auto &o : ctx->req.options) {
		if (o.name && !strcmp(o.name, name)) {
			iw_free(ctx, o.val);
			o.val = iw_strdup(ctx, val);
			return;
		}
	}
	if (ctx->req.options.size() >= 32) return;
	iwx_option o;
	o.name = iw_strdup(ctx, name);
	o.val = iw_strdup(ctx, val);
	ctx->req.options.push_back(o);
}

extern "C" const char *iw_get_option(struct iw_context *ctx, const char *name)
{
	for (auto &o : ctx->req.options)
		if (o.name && !strcmp(o.name, name)) return o.val;
	return nullptr;
}

// ---- utility functions (src/imagew-util.c:245-769) ------------------------

extern "C" int iw_file_to_memory(struct iw_context *ctx, struct iw_iodescr *io, void **pmem,
	iw_int64 *psize)
{
	*pmem = nullptr;
	*psize = 0;
	if (!io->getfilesize_fn) return 0;
	if (!io->getfilesize_fn(ctx, io, psize)) return 0;
	*pmem = iw_malloc(ctx, (size_t)*psize);
	size_t got = 0;
	if (!io->read_fn(ctx, io, *pmem, (size_t)*psize, &got)) return 0;
	return (iw_int64)got == *psize;
}

// UTF-8 to ASCII with a few substitutions (src/imagew-util.c:268-365).
extern "C" void iw_utf8_to_ascii(const char *src, char *dst, int dstlen)
{
	int dp = 0;
	auto emit = [&](unsigned char c) {
		if (dp > dstlen - 2) return;
		dst[dp++] = (char)c;
	};
	auto emit_uni = [&](unsigned int c) {
		const char *s = "?";
		switch (c) {
		case 0x00a9: s = "(c)"; break;
		case 0x00d7: s = "x"; break;
		case 0x2013: s = "-"; break;
		case 0x2018: case 0x2019: s = "'"; break;
		case 0x201c: case 0x201d: s = "\""; break;
		case 0x2192: s = "->"; break;
		case 0xfeff: s = ""; break;
		}
		for (; *s; s++) emit((unsigned char)*s);
	};
	unsigned int pending = 0;
	int expected = 0;
	for (int sp = 0; src[sp]; sp++) {
		unsigned char c = (unsigned char)src[sp];
		if (c < 128) {
			emit(c);
			expected = 0;
		} else if (c < 0xc0) {
			if (expected > 0) {
				pending = (pending << 6) | (c & 0x3f);
				if (--expected < 1) emit_uni(pending);
			}
		} else if (c < 0xe0) {
			pending = c & 0x1f;
			expected = 1;
		} else if (c < 0xf0) {
			pending = c & 0x0f;
			expected = 2;
		} else if (c < 0xf8) {
			pending = c & 0x07;
			expected = 3;
		}
	}
	dst[dp] = '\0';
}

extern "C" int iw_get_host_endianness(void)
{
	unsigned int x = 1;
	iw_byte b;
	memcpy(&b, &x, 1);
	return b == 0 ? IW_ENDIAN_BIG : IW_ENDIAN_LITTLE;
}

extern "C" void iw_set_ui16le(iw_byte *b, unsigned int n) { b[0] = n & 0xff; b[1] = (n >> 8) & 0xff; }
extern "C" void iw_set_ui32le(iw_byte *b, unsigned int n)
{
	for (int i = 0; i < 4; i++) b[i] = (n >> (8 * i)) & 0xff;
}
extern "C" void iw_set_ui16be(iw_byte *b, unsigned int n) { b[0] = (n >> 8) & 0xff; b[1] = n & 0xff; }
extern "C" void iw_set_ui32be(iw_byte *b, unsigned int n)
{
	for (int i = 0; i < 4; i++) b[i] = (n >> (24 - 8 * i)) & 0xff;
}
extern "C" unsigned int iw_get_ui16le(const iw_byte *b) { return (unsigned int)b[0] | ((unsigned int)b[1] << 8); }
extern "C" unsigned int iw_get_ui32le(const iw_byte *b)
{
	return (unsigned int)b[0] | ((unsigned int)b[1] << 8) | ((unsigned int)b[2] << 16) |
		((unsigned int)b[3] << 24);
}
extern "C" int iw_get_i32le(const iw_byte *b) { return (iw_int32)iw_get_ui32le(b); }
extern "C" unsigned int iw_get_ui16be(const iw_byte *b) { return ((unsigned int)b[0] << 8) | (unsigned int)b[1]; }
extern "C" unsigned int iw_get_ui32be(const iw_byte *b)
{
	return ((unsigned int)b[0] << 24) | ((unsigned int)b[1] << 16) | ((unsigned int)b[2] << 8) |
		(unsigned int)b[3];
}
extern "C" unsigned int iw_get_ui16_e(const iw_byte *b, int endian)
{
	return endian == IW_ENDIAN_LITTLE ? iw_get_ui16le(b) : iw_get_ui16be(b);
}
extern "C" unsigned int iw_get_ui32_e(const iw_byte *b, int endian)
{
	return endian == IW_ENDIAN_LITTLE ? iw_get_ui32le(b) : iw_get_ui32be(b);
}

extern "C" int iw_max_color_to_bitdepth(unsigned int mc)
{
	for (unsigned int bd = 1; bd <= 15; bd++)
		if (mc < (1U << bd)) return (int)bd;
	return 16;
}

// Format tables (src/imagew-util.c:456-675).
extern "C" int iw_detect_fmt_from_filename(const char *fn)
{
	static const struct { const char *ext; int fmt; } t[] = {
		{"png", IW_FORMAT_PNG}, {"jpg", IW_FORMAT_JPEG}, {"jpeg", IW_FORMAT_JPEG},
		{"bmp", IW_FORMAT_BMP}, {"tif", IW_FORMAT_TIFF}, {"tiff", IW_FORMAT_TIFF},
		{"miff", IW_FORMAT_MIFF}, {"webp", IW_FORMAT_WEBP}, {"gif", IW_FORMAT_GIF},
		{"pnm", IW_FORMAT_PNM}, {"pbm", IW_FORMAT_PBM}, {"pgm", IW_FORMAT_PGM},
		{"ppm", IW_FORMAT_PPM}, {"pam", IW_FORMAT_PAM}};
	const char *s = strrchr(fn, '.');
	if (!s) return IW_FORMAT_UNKNOWN;
	s++;
	for (const auto &e : t)
		if (!iw_stricmp(s, e.ext)) return e.fmt;
	return IW_FORMAT_UNKNOWN;
}

extern "C" const char *iw_get_fmt_name(int fmt)
{
	switch (fmt) {
	case IW_FORMAT_PNG: return "PNG";
	case IW_FORMAT_JPEG: return "JPEG";
	case IW_FORMAT_BMP: return "BMP";
	case IW_FORMAT_TIFF: return "TIFF";
	case IW_FORMAT_MIFF: return "MIFF";
	case IW_FORMAT_WEBP: return "WebP";
	case IW_FORMAT_GIF: return "GIF";
	case IW_FORMAT_PNM: return "PNM";
	case IW_FORMAT_PBM: return "PBM";
	case IW_FORMAT_PGM: return "PGM";
	case IW_FORMAT_PPM: return "PPM";
	case IW_FORMAT_PAM: return "PAM";
	}
	return nullptr;
}

extern "C" int iw_detect_fmt_of_file(const iw_byte *b, size_t n)
{
	if (n < 2) return IW_FORMAT_UNKNOWN;
	if (b[0] == 0x89 && b[1] == 0x50) return IW_FORMAT_PNG;
	if (n >= 3 && b[0] == 'G' && b[1] == 'I' && b[2] == 'F') return IW_FORMAT_GIF;
	if (b[0] == 0xff && b[1] == 0xd8) return IW_FORMAT_JPEG;
	if (b[0] == 'B' && (b[1] == 'M' || b[1] == 'A')) return IW_FORMAT_BMP;
	if ((b[0] == 0x49 || b[0] == 0x4d) && b[1] == b[0]) return IW_FORMAT_TIFF;
	if (b[0] == 0x69 && b[1] == 0x64) return IW_FORMAT_MIFF;
	if (n >= 12 && !memcmp(b, "RIFF", 4) && !memcmp(b + 8, "WEBP", 4)) return IW_FORMAT_WEBP;
	if (b[0] == 'P' && b[1] >= '1' && b[1] <= '6') return IW_FORMAT_PNM;
	if (b[0] == 'P' && b[1] == '7' && n >= 3 && b[2] == 0x0a) return IW_FORMAT_PAM;
	return IW_FORMAT_UNKNOWN;
}

extern "C" unsigned int iw_get_profile_by_fmt(int fmt)
{
	switch (fmt) {
	case IW_FORMAT_PNG:
		return IW_PROFILE_TRANSPARENCY | IW_PROFILE_GRAYSCALE | IW_PROFILE_PALETTETRNS |
			IW_PROFILE_GRAY1 | IW_PROFILE_GRAY2 | IW_PROFILE_GRAY4 | IW_PROFILE_16BPS |
			IW_PROFILE_BINARYTRNS | IW_PROFILE_PAL1 | IW_PROFILE_PAL2 | IW_PROFILE_PAL4 |
			IW_PROFILE_PAL8 | IW_PROFILE_REDUCEDBITDEPTHS | IW_PROFILE_PNG_BKGD;
	case IW_FORMAT_BMP:
		return IW_PROFILE_PAL1 | IW_PROFILE_PAL4 | IW_PROFILE_PAL8 | IW_PROFILE_REDUCEDBITDEPTHS;
	case IW_FORMAT_JPEG: return IW_PROFILE_GRAYSCALE;
	case IW_FORMAT_TIFF:
		return IW_PROFILE_TRANSPARENCY | IW_PROFILE_GRAYSCALE | IW_PROFILE_GRAY1 |
			IW_PROFILE_GRAY4 | IW_PROFILE_16BPS | IW_PROFILE_PAL4 | IW_PROFILE_PAL8;
	case IW_FORMAT_MIFF:
		return IW_PROFILE_TRANSPARENCY | IW_PROFILE_GRAYSCALE | IW_PROFILE_ALWAYSLINEAR |
			IW_PROFILE_HDRI | IW_PROFILE_RGB16_BKGD;
	case IW_FORMAT_WEBP: return 0;
	case IW_FORMAT_PAM:
		return IW_PROFILE_16BPS | IW_PROFILE_REDUCEDBITDEPTHS | IW_PROFILE_GRAYSCALE |
			IW_PROFILE_GRAY1 | IW_PROFILE_TRANSPARENCY;
	case IW_FORMAT_PNM:
		return IW_PROFILE_16BPS | IW_PROFILE_REDUCEDBITDEPTHS | IW_PROFILE_GRAYSCALE |
			IW_PROFILE_GRAY1;
	case IW_FORMAT_PPM: return IW_PROFILE_16BPS | IW_PROFILE_REDUCEDBITDEPTHS;
	case IW_FORMAT_PGM:
		return IW_PROFILE_16BPS | IW_PROFILE_REDUCEDBITDEPTHS | IW_PROFILE_GRAYSCALE;
	case IW_FORMAT_PBM: return IW_PROFILE_GRAY1;
	}
	return 0;
}

// The library reads and writes the dependency-free PNM/PAM family
// (iw_pnm.cpp); other file formats are outside the pixel path.
extern "C" int iw_is_input_fmt_supported(int fmt)
{
	return fmt == IW_FORMAT_PNM || fmt == IW_FORMAT_PAM;
}

extern "C" int iw_is_output_fmt_supported(int fmt)
{
	return fmt == IW_FORMAT_PNM || fmt == IW_FORMAT_PPM || fmt == IW_FORMAT_PGM ||
		fmt == IW_FORMAT_PBM || fmt == IW_FORMAT_PAM;
}

// Number parsing: comma-separated list, "a/b" rationals
// (src/imagew-util.c:677-769).
static int parse_one_number(const char *s, double *res, int *nread)
{
	int len = 0, slash = -1;
	*res = 0.0;
	*nread = 0;
	for (; s[len] && s[len] != ','; len++)
		if (s[len] == '/') slash = len;
	if (len < 1) return 0;
	*nread = len;
	if (slash >= 0) {
		double num = atof(s), den = atof(s + slash + 1);
		*res = (den == 0.0) ? 0.0 : num / den;
	} else {
		*res = atof(s);
	}
	return 1;
}

extern "C" int iw_parse_number_list(const char *s, int max_numbers, double *results)
{
	int pos = 0, count = 0;
	for (int n = 0; n < max_numbers; n++) {
		int nread;
		if (!parse_one_number(&s[pos], &results[n], &nread)) break;
		count++;
		pos += nread;
		if (s[pos] != ',') break;
		pos++;
	}
	return count;
}

extern "C" double iw_parse_number(const char *s)
{
	double r;
	int n;
	parse_one_number(s, &r, &n);
	return r;
}

extern "C" int iw_round_to_int(double x)
{
	if (x < 0.0) return -(int)(0.5 - x);
	return (int)(0.5 + x);
}

extern "C" int iw_parse_int(const char *s) { return iw_round_to_int(iw_parse_number(s)); }

// ---- codec entry points: not part of this library ------------------------

static int no_codec(iw_context *ctx)
{
	iw_set_error(ctx, "File formats are not supported by this library");
	return 0;
}

static char *no_version(char *s, int s_len)
{
	if (s && s_len > 0) s[0] = '\0';
	return s;
}

#define IWX_NO_CODEC(name) \
	extern "C" int name(struct iw_context *ctx, struct iw_iodescr *) { return no_codec(ctx); }
IWX_NO_CODEC(iw_read_png_file)
IWX_NO_CODEC(iw_write_png_file)
IWX_NO_CODEC(iw_read_jpeg_file)
IWX_NO_CODEC(iw_write_jpeg_file)
IWX_NO_CODEC(iw_read_bmp_file)
IWX_NO_CODEC(iw_write_bmp_file)
IWX_NO_CODEC(iw_write_tiff_file)
IWX_NO_CODEC(iw_read_webp_file)
IWX_NO_CODEC(iw_write_webp_file)
IWX_NO_CODEC(iw_read_gif_file)
extern "C" void iw_enable_zlib(struct iw_context *) {}
extern "C" char *iw_get_libpng_version_string(char *s, int n) { return no_version(s, n); }
extern "C" char *iw_get_zlib_version_string(char *s, int n) { return no_version(s, n); }
extern "C" char *iw_get_libjpeg_version_string(char *s, int n) { return no_version(s, n); }
extern "C" char *iw_get_libwebp_dec_version_string(char *s, int n) { return no_version(s, n); }
extern "C" char *iw_get_libwebp_enc_version_string(char *s, int n) { return no_version(s, n); }

// ---- library extensions ---------------------------------------------------

extern "C" void iwamd_set_precision(struct iw_context *ctx, int mode) { ctx->precision = mode; }

extern "C" size_t iwamd_plan_size(void) { return sizeof(iwamd_plan); }

extern "C" const char *iwamd_version(void) { return "imageworsener-amd 0.1 (gfx950)"; }

extern "C" int iwamd_device_count(void) { return iwx_gpu_device_count(); }

extern "C" int iwamd_last_timings(struct iw_context *ctx, double *ms, int n)
{
	for (int i = 0; i < n && i < ctx->num_last_timings; i++) ms[i] = ctx->last_timings_ms[i];
	return ctx->num_last_timings;
}
