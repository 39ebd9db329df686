// Host-side decisions taken before any pixel is touched, and the top-level
// iw_process_image().
//
// Every decision of the reference's iw_prepare_processing()
// (src/imagew-main.c:1899-2248) is taken here, on the host, in the reference's
// order, because later steps read the state earlier steps wrote.  The result
// is summarised in an iwamd_plan (include/iwamd.h) that the GPU executor
// (iw_gpu.hip) turns into device tables and kernel launches.
#include "iw_ctx.h"

#include <cmath>
#include <cstring>
#include <ctime>

// ---- colour transfer functions (src/imagew-main.c:28-96, 427-446) --------
// These run on the host only: they call libm pow(), whose last bit the
// output must reproduce.

double iwx_x_to_linear(double v, const struct iw_csdescr *cs)
{
	switch (cs->cstype) {
	case IW_CSTYPE_LINEAR:
		return v;
	case IW_CSTYPE_GAMMA:
		return pow(v, cs->gamma);
	case IW_CSTYPE_REC709:
		if (v < 4.5 * 0.020) return v / 4.5;
		return pow((v + 0.099) / 1.099, 1.0 / 0.45);
	default:  // sRGB, also the fallback
		if (v <= 0.04045) return v / 12.92;
		return pow((v + 0.055) / (1.055), 2.4);
	}
}

double iwx_linear_to_x(double v, const struct iw_csdescr *cs)
{
	if (v > 0.999999999) return 1.0;
	switch (cs->cstype) {
	case IW_CSTYPE_LINEAR:
		return v;
	case IW_CSTYPE_GAMMA:
		return pow(v, 1.0 / cs->gamma);
	case IW_CSTYPE_REC709:
		if (v < 0.020) return 4.5 * v;
		return 1.099 * pow(v, 0.45) - 0.099;
	default:
		if (v <= 0.0031308) return 12.92 * v;
		return 1.055 * pow(v, 1.0 / 2.4) - 0.055;
	}
}

extern "C" double iw_convert_sample_to_linear(double v, const struct iw_csdescr *cs)
{
	return iwx_x_to_linear(v, cs);
}

extern "C" double iw_convert_sample_from_linear(double v, const struct iw_csdescr *cs)
{
	return iwx_linear_to_x(v, cs);
}

// Weighted sum ((w0*r + w1*g) + w2*b), optionally of the sorted components
// (src/imagew-main.c:278-305).
double iwx_color_to_grayscale(iw_context *ctx, double r, double g, double b)
{
	const double *w = ctx->grayscale_weight;
	if (ctx->grayscale_formula == IW_GSF_WEIGHTED) return w[0] * r + w[1] * g + w[2] * b;
	if (ctx->grayscale_formula == IW_GSF_ORDERBYVALUE) {
		double hi, mid, lo;
		if (g <= r) { hi = r; mid = g; } else { hi = g; mid = r; }
		if (b <= mid) {
			lo = b;
		} else {
			lo = mid;
			if (b <= hi) mid = b;
			else { mid = hi; hi = b; }
		}
		return w[0] * hi + w[1] * mid + w[2] * lo;
	}
	return 0.0;
}

// ---- channel layout ------------------------------------------------------

static int alpha_index_of(int t)
{
	if (t == IW_IMGTYPE_RGBA) return 3;
	if (t == IW_IMGTYPE_GRAYA) return 1;
	return 0;
}

static int channeltype_of(int imgtype, int ch)
{
	static const int gray[] = {IW_CHANNELTYPE_GRAY, IW_CHANNELTYPE_ALPHA};
	static const int rgb[] = {IW_CHANNELTYPE_RED, IW_CHANNELTYPE_GREEN, IW_CHANNELTYPE_BLUE,
		IW_CHANNELTYPE_ALPHA};
	switch (imgtype) {
	case IW_IMGTYPE_GRAY: return ch == 0 ? gray[0] : 0;
	case IW_IMGTYPE_GRAYA: return ch < 2 ? gray[ch] : 0;
	case IW_IMGTYPE_RGB: return ch < 3 ? rgb[ch] : 0;
	case IW_IMGTYPE_RGBA: return ch < 4 ? rgb[ch] : 0;
	}
	return 0;
}

// Logical input type gains a virtual alpha channel when either dimension uses
// TRANSPARENT edges (src/imagew-main.c:1842-1880).
static void init_channel_info(iw_context *ctx)
{
	ctx->img1_imgtype_logical = ctx->img1.imgtype;
	if (ctx->resize_settings[IW_DIMENSION_H].edge_policy == IW_EDGE_POLICY_TRANSPARENT ||
		ctx->resize_settings[IW_DIMENSION_V].edge_policy == IW_EDGE_POLICY_TRANSPARENT) {
		if (ctx->img1.imgtype == IW_IMGTYPE_GRAY) ctx->img1_imgtype_logical = IW_IMGTYPE_GRAYA;
		else if (ctx->img1.imgtype == IW_IMGTYPE_RGB) ctx->img1_imgtype_logical = IW_IMGTYPE_RGBA;
	}
	ctx->img1_numchannels_physical = iw_imgtype_num_channels(ctx->img1.imgtype);
	ctx->img1_numchannels_logical = iw_imgtype_num_channels(ctx->img1_imgtype_logical);
	ctx->img1_alpha_channel_index = alpha_index_of(ctx->img1_imgtype_logical);

	for (int i = 0; i < ctx->img1_numchannels_logical; i++)
		ctx->img1_ci[i].channeltype = channeltype_of(ctx->img1_imgtype_logical, i);

	ctx->img2.imgtype = ctx->img1_imgtype_logical;
	ctx->img2_numchannels = ctx->img1_numchannels_logical;
	ctx->intermed_numchannels = ctx->img1_numchannels_logical;

	for (int i = 0; i < ctx->img1_numchannels_logical; i++) {
		ctx->intermed_ci[i].channeltype = ctx->img1_ci[i].channeltype;
		ctx->intermed_ci[i].corresponding_input_channel = i;
		ctx->img2_ci[i].channeltype = ctx->img1_ci[i].channeltype;
		if (i >= ctx->img1_numchannels_physical) ctx->img1_ci[i].disable_fast_get_sample = 1;
	}
}

// ---- background ----------------------------------------------------------

static int bkgd_has_transparency(iw_context *ctx)
{
	if (!ctx->apply_bkgd) return 0;
	if (!(ctx->output_profile & IW_PROFILE_TRANSPARENCY)) return 0;
	if (ctx->apply_bkgd_strategy == IWX_BKGD_EARLY) return 0;
	if (ctx->bkgd_color_source == IWX_BKGD_SRC_FILE)
		return ctx->img1_bkgd_label_inputcs.c[3] < 1.0;
	if (ctx->bkgd_color_source == IWX_BKGD_SRC_REQ) {
		if (ctx->bkgd_checkerboard && ctx->req.bkgd2.c[3] < 1.0) return 1;
		if (ctx->req.bkgd.c[3] < 1.0) return 1;
	}
	return 0;
}

// When and from where a background is applied (src/imagew-main.c:1734-1823).
static void decide_how_to_apply_bkgd(iw_context *ctx)
{
	if (!IW_IMGTYPE_HAS_ALPHA(ctx->img1_imgtype_logical)) {
		ctx->apply_bkgd = 0;
		return;
	}
	if (ctx->img1_bkgd_label_set && (ctx->req.use_bkgd_label_from_file || !ctx->req.bkgd_valid)) {
		ctx->bkgd_color_source = IWX_BKGD_SRC_FILE;
	} else if (ctx->req.bkgd_valid) {
		ctx->bkgd_color_source = IWX_BKGD_SRC_REQ;
		ctx->bkgd_checkerboard = ctx->req.bkgd_checkerboard;
	} else {
		ctx->bkgd_color_source = IWX_BKGD_SRC_NONE;
	}
	if (ctx->bkgd_checkerboard && ctx->bkgd_check_size < 1) ctx->bkgd_check_size = 1;
	if (ctx->req.bkgd_valid) ctx->apply_bkgd = 1;

	if (!(ctx->output_profile & IW_PROFILE_TRANSPARENCY)) {
		if (!ctx->req.bkgd_valid && !ctx->apply_bkgd)
			iw_warning(ctx, "This image may have transparency, which is incompatible with the output format. A background color will be applied.");
		ctx->apply_bkgd = 1;
	}

	if (ctx->resize_settings[IW_DIMENSION_H].use_offset ||
		ctx->resize_settings[IW_DIMENSION_V].use_offset) {
		if (!ctx->req.bkgd_valid && !ctx->apply_bkgd)
			iw_warning(ctx, "This image may have transparency, which is incompatible with a channel offset. A background color will be applied.");
		ctx->apply_bkgd = 1;
		if (ctx->bkgd_checkerboard && ctx->req.bkgd_checkerboard) {
			iw_warning(ctx, "Checkerboard backgrounds are not supported when using a channel offset.");
			ctx->bkgd_checkerboard = 0;
		}
		ctx->apply_bkgd_strategy = IWX_BKGD_EARLY;
		return;
	}
	if (!ctx->apply_bkgd) return;
	// Checkerboards need LATE; solid colours also use LATE (the safe choice).
	ctx->apply_bkgd_strategy = IWX_BKGD_LATE;
}

// Background colours per channel, in linear light (src/imagew-main.c:1546-1611).
static void prepare_apply_bkgd(iw_context *ctx)
{
	struct iw_color b1 = {{1.0, 0.0, 1.0, 1.0}};
	struct iw_color b2 = {{0.0, 0.0, 0.0, 1.0}};
	if (!ctx->apply_bkgd) return;
	if (ctx->bkgd_color_source == IWX_BKGD_SRC_FILE) {
		b1 = ctx->img1_bkgd_label_lin;
		ctx->bkgd_checkerboard = 0;
	} else if (ctx->bkgd_color_source == IWX_BKGD_SRC_REQ) {
		b1 = ctx->req.bkgd;
		if (ctx->req.bkgd_checkerboard) b2 = ctx->req.bkgd2;
	}
	if (ctx->apply_bkgd_strategy == IWX_BKGD_EARLY) {
		ctx->bkgd1alpha = 1.0;
	} else {
		ctx->bkgd1alpha = b1.c[3];
		ctx->bkgd2alpha = b2.c[3];
	}
	const int out_rgb = ctx->img2.imgtype == IW_IMGTYPE_RGB || ctx->img2.imgtype == IW_IMGTYPE_RGBA;
	const int out_gray = ctx->img2.imgtype == IW_IMGTYPE_GRAY || ctx->img2.imgtype == IW_IMGTYPE_GRAYA;
	if (ctx->apply_bkgd_strategy == IWX_BKGD_LATE && out_rgb) {
		for (int i = 0; i < 3; i++) ctx->img2_ci[i].bkgd1_color_lin = b1.c[i];
		if (ctx->bkgd_checkerboard)
			for (int i = 0; i < 3; i++) ctx->img2_ci[i].bkgd2_color_lin = b2.c[i];
	} else if (ctx->apply_bkgd_strategy == IWX_BKGD_LATE && out_gray) {
		ctx->img2_ci[0].bkgd1_color_lin = iwx_color_to_grayscale(ctx, b1.c[0], b1.c[1], b1.c[2]);
		if (ctx->bkgd_checkerboard)
			ctx->img2_ci[0].bkgd2_color_lin = iwx_color_to_grayscale(ctx, b2.c[0], b2.c[1], b2.c[2]);
	} else if (ctx->apply_bkgd_strategy == IWX_BKGD_EARLY && ctx->img2.imgtype == IW_IMGTYPE_RGB) {
		for (int i = 0; i < 3; i++) ctx->intermed_ci[i].bkgd_color_lin = b1.c[i];
	} else if (ctx->apply_bkgd_strategy == IWX_BKGD_EARLY && ctx->img2.imgtype == IW_IMGTYPE_GRAY) {
		ctx->intermed_ci[0].bkgd_color_lin = iwx_color_to_grayscale(ctx, b1.c[0], b1.c[1], b1.c[2]);
	}
}

// ---- strategy: input -> intermediate -> output types
// (src/imagew-main.c:1613-1718, 2049-2131) --------------------------------
//
// A strategy is encoded as (from_type << 4 | to_type) with 1=G 2=GA 3=RGB
// 4=RGBA; stage 1 maps input to intermediate, stage 2 intermediate to output.
enum { T_G = 1, T_GA = 2, T_RGB = 3, T_RGBA = 4 };
#define STRAT(a, b) (((a) << 4) | (b))

static void decide_strategy(iw_context *ctx, int *ps1, int *ps2)
{
	int s1, s2;
	const int gs = ctx->to_grayscale;
	switch (ctx->img1_imgtype_logical) {
	case IW_IMGTYPE_RGBA:
		s1 = gs ? STRAT(T_RGBA, T_GA) : STRAT(T_RGBA, T_RGBA);
		s2 = gs ? STRAT(T_GA, T_GA) : STRAT(T_RGBA, T_RGBA);
		break;
	case IW_IMGTYPE_RGB:
		s1 = gs ? STRAT(T_RGB, T_G) : STRAT(T_RGB, T_RGB);
		s2 = gs ? STRAT(T_G, T_G) : STRAT(T_RGB, T_RGB);
		break;
	case IW_IMGTYPE_GRAYA:
		s1 = gs ? STRAT(T_GA, T_GA) : STRAT(T_GA, T_RGBA);
		s2 = gs ? STRAT(T_GA, T_GA) : STRAT(T_RGBA, T_RGBA);
		break;
	default:
		s1 = gs ? STRAT(T_G, T_G) : STRAT(T_G, T_RGB);
		s2 = gs ? STRAT(T_G, T_G) : STRAT(T_RGB, T_RGB);
	}
	if (ctx->apply_bkgd && ctx->apply_bkgd_strategy == IWX_BKGD_EARLY) {
		// Background applied before resizing: the intermediate loses alpha.
		if (s1 == STRAT(T_RGBA, T_RGBA)) { s1 = STRAT(T_RGBA, T_RGB); s2 = STRAT(T_RGB, T_RGB); }
		else if (s1 == STRAT(T_GA, T_GA)) { s1 = STRAT(T_GA, T_G); s2 = STRAT(T_G, T_G); }
		else if (s1 == STRAT(T_GA, T_RGBA)) { s1 = STRAT(T_GA, T_RGB); s2 = STRAT(T_RGB, T_RGB); }
		else if (s1 == STRAT(T_RGBA, T_GA)) { s1 = STRAT(T_RGBA, T_G); s2 = STRAT(T_G, T_G); }
	}
	if (ctx->apply_bkgd && !bkgd_has_transparency(ctx)) {
		// Opaque background applied after resizing: the output loses alpha.
		if (s2 == STRAT(T_GA, T_GA)) s2 = STRAT(T_GA, T_G);
		else if (s2 == STRAT(T_RGBA, T_RGBA)) s2 = STRAT(T_RGBA, T_RGB);
	}
	*ps1 = s1;
	*ps2 = s2;
}

static int apply_strategy(iw_context *ctx, int s1, int s2)
{
	iwx_ci_intermed *ic = ctx->intermed_ci;
	switch (s1) {
	case STRAT(T_RGBA, T_RGBA): ctx->intermed_imgtype = IW_IMGTYPE_RGBA; break;
	case STRAT(T_GA, T_RGBA):
		ctx->intermed_imgtype = IW_IMGTYPE_RGBA;
		ic[0].corresponding_input_channel = ic[1].corresponding_input_channel =
			ic[2].corresponding_input_channel = 0;
		ic[3].corresponding_input_channel = 1;
		break;
	case STRAT(T_RGB, T_RGB):
	case STRAT(T_RGBA, T_RGB): ctx->intermed_imgtype = IW_IMGTYPE_RGB; break;
	case STRAT(T_G, T_RGB):
	case STRAT(T_GA, T_RGB):
		ctx->intermed_imgtype = IW_IMGTYPE_RGB;
		ic[0].corresponding_input_channel = ic[1].corresponding_input_channel =
			ic[2].corresponding_input_channel = 0;
		break;
	case STRAT(T_RGBA, T_GA):
		ctx->intermed_imgtype = IW_IMGTYPE_GRAYA;
		ic[0].cvt_to_grayscale = 1;
		ic[0].corresponding_input_channel = 0;
		ic[1].corresponding_input_channel = 3;
		break;
	case STRAT(T_GA, T_GA): ctx->intermed_imgtype = IW_IMGTYPE_GRAYA; break;
	case STRAT(T_RGB, T_G):
		ctx->intermed_imgtype = IW_IMGTYPE_GRAY;
		ic[0].cvt_to_grayscale = 1;
		ic[0].corresponding_input_channel = 0;
		break;
	case STRAT(T_G, T_G):
		ctx->intermed_imgtype = IW_IMGTYPE_GRAY;
		ic[0].corresponding_input_channel = 0;
		break;
	default:
		// GA->G and RGBA->G exist in the reference's table but are unreachable;
		// it reports them as an internal error.
		iw_set_errorf(ctx, "Internal error, unknown strategy %d", s1);
		return 0;
	}
	ctx->intermed_numchannels = iw_imgtype_num_channels(ctx->intermed_imgtype);
	ctx->intermed_alpha_channel_index = alpha_index_of(ctx->intermed_imgtype);
	for (int i = 0; i < ctx->intermed_numchannels; i++) ic[i].corresponding_output_channel = i;

	switch (s2) {
	case STRAT(T_RGBA, T_RGBA): ctx->img2.imgtype = IW_IMGTYPE_RGBA; break;
	case STRAT(T_RGB, T_RGB): ctx->img2.imgtype = IW_IMGTYPE_RGB; break;
	case STRAT(T_RGBA, T_RGB):
		ctx->img2.imgtype = IW_IMGTYPE_RGB;
		ic[3].corresponding_output_channel = -1;
		break;
	case STRAT(T_GA, T_GA): ctx->img2.imgtype = IW_IMGTYPE_GRAYA; break;
	case STRAT(T_G, T_G): ctx->img2.imgtype = IW_IMGTYPE_GRAY; break;
	case STRAT(T_GA, T_G):
		ctx->img2.imgtype = IW_IMGTYPE_GRAY;
		ic[1].corresponding_output_channel = -1;
		break;
	default:
		iw_set_error(ctx, "Internal error");
		return 0;
	}
	ctx->img2_numchannels = iw_imgtype_num_channels(ctx->img2.imgtype);
	for (int i = 0; i < ctx->intermed_numchannels; i++)
		ic[i].channeltype = channeltype_of(ctx->intermed_imgtype, i);
	for (int i = 0; i < ctx->img2_numchannels; i++)
		ctx->img2_ci[i].channeltype = channeltype_of(ctx->img2.imgtype, i);
	return 1;
}

// Output sample type and depth (src/imagew-main.c:1514-1541).
static void decide_output_bit_depth(iw_context *ctx)
{
	if (ctx->output_profile & IW_PROFILE_HDRI) {
		ctx->img2.sampletype = IW_SAMPLETYPE_FLOATINGPOINT;
		ctx->img2.bit_depth = 32;
		return;
	}
	ctx->img2.sampletype = IW_SAMPLETYPE_UINT;
	if (ctx->req.output_depth > 8 && (ctx->output_profile & IW_PROFILE_16BPS)) {
		ctx->img2.bit_depth = 16;
	} else {
		if (ctx->req.output_depth > 8)
			iw_warning(ctx, "Reducing depth to 8; required by the output format.");
		ctx->img2.bit_depth = 8;
	}
}

// NULL when the size along a dimension is unchanged (and nothing moves it),
// else Catmull-Rom (src/imagew-main.c:1825-1840).
static void set_auto_resizetype(iw_context *ctx, int size1, int size2, int dim)
{
	const iwx_resize_settings &rs = ctx->resize_settings[dim];
	if (size2 == size1 && !rs.use_offset && !ctx->req.out_true_valid && rs.translate == 0.0)
		iw_set_resize_alg(ctx, dim, IW_RESIZETYPE_NULL, 1.0, 0.0, 0.0);
	else
		iw_set_resize_alg(ctx, dim, IW_RESIZETYPE_CUBIC, 1.0, 0.0, 0.5);
}

int iwx_prepare_processing(iw_context *ctx, int w, int h)
{
	if (ctx->output_profile == 0) {
		iw_set_error(ctx, "Output profile not set");
		return 0;
	}
	if (ctx->randomize) ctx->random_seed = (int)time(nullptr);

	if (ctx->req.out_true_valid) {
		ctx->resize_settings[IW_DIMENSION_H].out_true_size = ctx->req.out_true_width;
		ctx->resize_settings[IW_DIMENSION_V].out_true_size = ctx->req.out_true_height;
	} else {
		ctx->resize_settings[IW_DIMENSION_H].out_true_size = (double)w;
		ctx->resize_settings[IW_DIMENSION_V].out_true_size = (double)h;
	}
	if (!iw_check_image_dimensions(ctx, ctx->img1.width, ctx->img1.height)) return 0;
	if (!iw_check_image_dimensions(ctx, w, h)) return 0;

	if (ctx->to_grayscale) {
		if (ctx->grayscale_formula == IW_GSF_STANDARD) {
			ctx->grayscale_formula = IW_GSF_WEIGHTED;
			iw_set_grayscale_weights(ctx, 0.212655, 0.715158, 0.072187);
		} else if (ctx->grayscale_formula == IW_GSF_COMPATIBLE) {
			ctx->grayscale_formula = IW_GSF_WEIGHTED;
			iw_set_grayscale_weights(ctx, 0.299, 0.587, 0.114);
		}
	}

	init_channel_info(ctx);
	ctx->img2.width = w;
	ctx->img2.height = h;

	// Source region.  The y clamp assigns input_start_x, as the reference
	// does (src/imagew-main.c:1952); callers rely on identical results.
	if (ctx->input_start_x < 0) ctx->input_start_x = 0;
	if (ctx->input_start_y < 0) ctx->input_start_y = 0;
	if (ctx->input_start_x > ctx->img1.width - 1) ctx->input_start_x = ctx->img1.width - 1;
	if (ctx->input_start_y > ctx->img1.height - 1) ctx->input_start_x = ctx->img1.height - 1;
	if (ctx->input_w < 0) ctx->input_w = ctx->img1.width - ctx->input_start_x;
	if (ctx->input_h < 0) ctx->input_h = ctx->img1.height - ctx->input_start_y;
	if (ctx->input_w < 1) ctx->input_w = 1;
	if (ctx->input_h < 1) ctx->input_h = 1;
	if (ctx->input_w > ctx->img1.width - ctx->input_start_x)
		ctx->input_w = ctx->img1.width - ctx->input_start_x;
	if (ctx->input_h > ctx->img1.height - ctx->input_start_y)
		ctx->input_h = ctx->img1.height - ctx->input_start_y;

	// Output colour space.
	if (ctx->req.output_cs_valid) {
		ctx->img2cs = ctx->req.output_cs;
		if ((ctx->output_profile & IW_PROFILE_ALWAYSLINEAR) && ctx->img2cs.cstype != IW_CSTYPE_LINEAR) {
			iw_warning(ctx, "Forcing output colorspace to linear; required by the output format.");
			iw_make_linear_csdescr(&ctx->img2cs);
		}
	} else if (ctx->output_profile & IW_PROFILE_ALWAYSLINEAR) {
		iw_make_linear_csdescr(&ctx->img2cs);
	} else {
		iw_make_srgb_csdescr_2(&ctx->img2cs);
	}

	// Input max colour codes; any reduced code forces the slow sample path.
	if (ctx->img1.sampletype != IW_SAMPLETYPE_FLOATINGPOINT) {
		ctx->input_maxcolorcode_int = (1 << ctx->img1.bit_depth) - 1;
		ctx->input_maxcolorcode = (double)ctx->input_maxcolorcode_int;
		for (int i = 0; i < IWX_CI_COUNT; i++) {
			if (ctx->img1_ci[i].maxcolorcode_int <= 0)
				ctx->img1_ci[i].maxcolorcode_int = ctx->input_maxcolorcode_int;
			ctx->img1_ci[i].maxcolorcode_dbl = (double)ctx->img1_ci[i].maxcolorcode_int;
			if (ctx->img1_ci[i].maxcolorcode_int != ctx->input_maxcolorcode_int)
				ctx->support_reduced_input_bitdepths = 1;
		}
	}
	if (ctx->support_reduced_input_bitdepths || ctx->img1.sampletype == IW_SAMPLETYPE_FLOATINGPOINT)
		for (int i = 0; i < ctx->img1_numchannels_physical; i++) ctx->img1_ci[i].disable_fast_get_sample = 1;

	for (int d = 0; d < 2; d++)
		for (int c = 0; c < 3; c++)
			if (fabs(ctx->resize_settings[d].channel_offset[c]) > 0.00001) ctx->resize_settings[d].use_offset = 1;
	if (ctx->to_grayscale && (ctx->resize_settings[0].use_offset || ctx->resize_settings[1].use_offset)) {
		iw_warning(ctx, "Disabling channel offset, due to grayscale output.");
		ctx->resize_settings[0].use_offset = 0;
		ctx->resize_settings[1].use_offset = 0;
	}

	decide_how_to_apply_bkgd(ctx);

	// Weight lists differ per channel with channel offsets, or when virtual
	// pixels take the (per-channel) background colour.
	for (int d = 0; d < 2; d++) {
		if (ctx->resize_settings[d].use_offset ||
			(ctx->apply_bkgd && ctx->apply_bkgd_strategy == IWX_BKGD_EARLY &&
			 ctx->resize_settings[d].edge_policy == IW_EDGE_POLICY_TRANSPARENT))
			ctx->disable_rrctx_cache[d] = 1;
	}

	int s1, s2;
	decide_strategy(ctx, &s1, &s2);
	if (!apply_strategy(ctx, s1, s2)) return 0;

	if (IW_IMGTYPE_HAS_ALPHA(ctx->intermed_imgtype))
		for (int i = 0; i < ctx->intermed_numchannels; i++)
			if (ctx->intermed_ci[i].channeltype != IW_CHANNELTYPE_ALPHA)
				ctx->intermed_ci[i].need_unassoc_alpha_processing = 1;

	decide_output_bit_depth(ctx);

	if (ctx->img2.sampletype == IW_SAMPLETYPE_FLOATINGPOINT) {
		int any = 0;
		for (int i = 0; i < IWX_NUM_CHANNELTYPES; i++) any |= ctx->req.color_count[i] != 0;
		if (any) iw_warning(ctx, "Posterization is not supported with floating point output.");
	} else {
		const int full = (1 << ctx->img2.bit_depth) - 1;
		for (int i = 0; i < ctx->img2_numchannels; i++) ctx->img2_ci[i].maxcolorcode_int = full;
		if (ctx->output_profile & IW_PROFILE_REDUCEDBITDEPTHS) {
			for (int i = 0; i < ctx->img2_numchannels; i++) {
				int m = ctx->req.output_maxcolorcode[ctx->img2_ci[i].channeltype];
				if (m > 0) ctx->img2_ci[i].maxcolorcode_int = m > full ? full : m;
			}
		}
		for (int i = 0; i < ctx->img2_numchannels; i++) {
			if (ctx->img2_ci[i].maxcolorcode_int != full) {
				ctx->reduced_output_maxcolor_flag = 1;
				ctx->disable_output_lookup_tables = 1;
			}
			ctx->img2_ci[i].maxcolorcode_dbl = (double)ctx->img2_ci[i].maxcolorcode_int;
		}
	}

	for (int i = 0; i < ctx->img2_numchannels; i++) {
		iwx_ci_out &oc = ctx->img2_ci[i];
		oc.color_count = ctx->req.color_count[oc.channeltype];
		if (oc.color_count) {
			if (oc.color_count < 2) oc.color_count = 2;
			else if (oc.color_count > oc.maxcolorcode_int) oc.color_count = oc.maxcolorcode_int;
		}
		if (oc.color_count == 1 + oc.maxcolorcode_int) oc.color_count = 0;
		oc.ditherfamily = ctx->ditherfamily_by_channeltype[oc.channeltype];
		oc.dithersubtype = ctx->dithersubtype_by_channeltype[oc.channeltype];
		if (oc.ditherfamily == IW_DITHERFAMILY_ERRDIFF) ctx->uses_errdiffdither = 1;
	}

	if (ctx->img1_bkgd_label_set) {
		for (int i = 0; i < 3; i++)
			ctx->img1_bkgd_label_lin.c[i] = iwx_x_to_linear(ctx->img1_bkgd_label_inputcs.c[i], &ctx->img1cs);
		ctx->img1_bkgd_label_lin.c[3] = ctx->img1_bkgd_label_inputcs.c[3];
	}
	if (ctx->apply_bkgd) prepare_apply_bkgd(ctx);

	ctx->img2.rendering_intent = (ctx->req.output_rendering_intent == IW_INTENT_UNKNOWN)
		? ctx->img1.rendering_intent : ctx->req.output_rendering_intent;

	if (ctx->resize_settings[IW_DIMENSION_H].family == IW_RESIZETYPE_AUTO)
		set_auto_resizetype(ctx, ctx->input_w, ctx->img2.width, IW_DIMENSION_H);
	if (ctx->resize_settings[IW_DIMENSION_V].family == IW_RESIZETYPE_AUTO)
		set_auto_resizetype(ctx, ctx->input_h, ctx->img2.height, IW_DIMENSION_V);

	if (IW_IMGTYPE_HAS_ALPHA(ctx->img2.imgtype) && !ctx->opt_strip_alpha) {
		ctx->opt_palette = 0;
		ctx->opt_binary_trns = 0;
	}

	// Buffers the reference allocates in iw_process_internal
	// (src/imagew-main.c:1380-1402); their limits still apply.
	ctx->intermed_canvas_width = ctx->input_w;
	ctx->intermed_canvas_height = ctx->img2.height;
	ctx->img2.bpr = iw_calc_bytesperrow(ctx->img2.width, ctx->img2.bit_depth * ctx->img2_numchannels);
	if (!iwx_alloc_check(ctx, ctx->img2.bpr, ctx->img2.height)) return 0;
	if (!iwx_alloc_check(ctx, (size_t)ctx->intermed_canvas_width * ctx->intermed_canvas_height, 4)) return 0;
	if (ctx->uses_errdiffdither && !iwx_alloc_check(ctx, (size_t)ctx->img2.width, 8)) return 0;

	// The nearest-colour table exists only for 8-bit (or smaller) non-linear
	// UINT output larger than 512 pixels (src/imagew-main.c:1197-1236, 1404-1408).
	ctx->nearest_color_table_exists = !ctx->disable_output_lookup_tables && !ctx->no_gamma &&
		ctx->img2cs.cstype != IW_CSTYPE_LINEAR &&
		ctx->img2.sampletype != IW_SAMPLETYPE_FLOATINGPOINT &&
		ctx->img2.bit_depth == ctx->img2.bit_depth && ctx->img2.bit_depth <= 8 &&
		((size_t)ctx->img2.width) * ctx->img2.height > 512;

	if (IW_IMGTYPE_HAS_ALPHA(ctx->intermed_imgtype)) {
		if (!iwx_alloc_check(ctx, (size_t)ctx->intermed_canvas_width * ctx->intermed_canvas_height, 4)) return 0;
		if (!iwx_alloc_check(ctx, (size_t)ctx->img2.width * ctx->img2.height, 4)) return 0;
	}
	return 1;
}

// ---- the plan ------------------------------------------------------------

void iwx_fill_plan(iw_context *ctx, iwamd_plan *p)
{
	memset(p, 0, sizeof(*p));
	struct iw_csdescr lin;
	iw_make_linear_csdescr(&lin);

	p->in_pixels = ctx->img1.pixels;
	p->in_bpr = ctx->img1.bpr;
	p->in_width = ctx->img1.width;
	p->in_height = ctx->img1.height;
	p->in_bit_depth = ctx->img1.bit_depth;
	p->in_sampletype = ctx->img1.sampletype;
	p->in_imgtype = ctx->img1.imgtype;
	p->in_nch_physical = ctx->img1_numchannels_physical;
	p->orient_transform = ctx->img1.orient_transform;
	p->input_start_x = ctx->input_start_x;
	p->input_start_y = ctx->input_start_y;
	p->input_w = ctx->input_w;
	p->input_h = ctx->input_h;
	p->in_alpha_index = ctx->img1_alpha_channel_index;
	p->in_has_alpha = IW_IMGTYPE_HAS_ALPHA(ctx->img1_imgtype_logical);
	p->in_maxcolorcode = ctx->input_maxcolorcode;
	for (int i = 0; i < IWAMD_MAX_CH; i++) {
		p->in_ch_maxcolorcode[i] = ctx->img1_ci[i].maxcolorcode_dbl;
		p->in_ch_slow[i] = ctx->img1_ci[i].disable_fast_get_sample;
	}
	p->grayscale_formula = ctx->grayscale_formula;
	for (int i = 0; i < 3; i++) p->grayscale_weight[i] = ctx->grayscale_weight[i];

	p->inter_nch = ctx->intermed_numchannels;
	p->inter_alpha_index = IW_IMGTYPE_HAS_ALPHA(ctx->intermed_imgtype) ? ctx->intermed_alpha_channel_index : -1;
	p->inter_width = ctx->intermed_canvas_width;
	p->inter_height = ctx->intermed_canvas_height;
	const int early = ctx->apply_bkgd && ctx->apply_bkgd_strategy == IWX_BKGD_EARLY;
	for (int i = 0; i < ctx->intermed_numchannels; i++) {
		const iwx_ci_intermed &s = ctx->intermed_ci[i];
		iwamd_ichan &d = p->ich[i];
		d.channeltype = s.channeltype;
		d.input_channel = s.corresponding_input_channel;
		d.cvt_to_grayscale = s.cvt_to_grayscale;
		d.output_channel = s.corresponding_output_channel;
		d.need_unassoc_alpha = s.need_unassoc_alpha_processing;
		d.early_bkgd = !s.need_unassoc_alpha_processing && early;
		d.bkgd_color_lin = s.bkgd_color_lin;
		const int is_alpha = s.channeltype == IW_CHANNELTYPE_ALPHA;
		d.in_cs = (is_alpha || ctx->no_gamma) ? lin : ctx->img1cs;
		for (int dim = 0; dim < 2; dim++) {
			// Virtual (TRANSPARENT) pixels take the channel's background colour
			// when it is applied before resizing (src/imagew-resize.c:526-532).
			d.edge_sample_value[dim] = (ctx->resize_settings[dim].edge_policy == IW_EDGE_POLICY_TRANSPARENT && early)
				? s.bkgd_color_lin : 0.0;
		}
	}
	// Without per-channel weight lists the reference reuses the list built for
	// the first channel it processes (alpha when present), including that
	// channel's virtual-pixel value.
	const int first = (p->inter_alpha_index >= 0) ? p->inter_alpha_index : 0;
	for (int dim = 0; dim < 2; dim++) {
		if (ctx->disable_rrctx_cache[dim]) continue;
		for (int i = 0; i < ctx->intermed_numchannels; i++)
			p->ich[i].edge_sample_value[dim] = p->ich[first].edge_sample_value[dim];
	}
	for (int dim = 0; dim < 2; dim++) {
		const iwx_resize_settings &s = ctx->resize_settings[dim];
		iwamd_resize &r = p->rs[dim];
		r.family = s.family;
		r.edge_policy = s.edge_policy;
		r.param1 = s.param1;
		r.param2 = s.param2;
		r.blur_factor = s.blur_factor;
		r.out_true_size = s.out_true_size;
		r.translate = s.translate;
		r.use_offset = s.use_offset;
		for (int c = 0; c < 3; c++) r.channel_offset[c] = s.channel_offset[c];
	}
	p->rs[IW_DIMENSION_V].num_in = ctx->input_h;
	p->rs[IW_DIMENSION_V].num_out = ctx->intermed_canvas_height;
	p->rs[IW_DIMENSION_H].num_in = ctx->intermed_canvas_width;
	p->rs[IW_DIMENSION_H].num_out = ctx->img2.width;
	p->intclamp = ctx->intclamp;

	p->out_width = ctx->img2.width;
	p->out_height = ctx->img2.height;
	p->out_bit_depth = ctx->img2.bit_depth;
	p->out_sampletype = ctx->img2.sampletype;
	p->out_imgtype = ctx->img2.imgtype;
	p->out_nch = ctx->img2_numchannels;
	p->out_bpr = ctx->img2.bpr;
	for (int o = 0; o < ctx->img2_numchannels; o++) {
		const iwx_ci_out &s = ctx->img2_ci[o];
		iwamd_ochan &d = p->och[o];
		d.channeltype = s.channeltype;
		d.ditherfamily = s.ditherfamily;
		d.dithersubtype = s.dithersubtype;
		d.color_count = s.color_count;
		d.maxcolorcode_int = s.maxcolorcode_int;
		d.maxcolorcode_dbl = s.maxcolorcode_dbl;
		const int is_alpha = s.channeltype == IW_CHANNELTYPE_ALPHA;
		d.use_nc_table = ctx->nearest_color_table_exists && !is_alpha &&
			s.ditherfamily == IW_DITHERFAMILY_NONE && s.color_count == 0;
		// Alpha has its own seed; "r" (not "r2") gives every channel its own
		// (src/imagew-main.c:1002-1012).
		d.random_seed = (s.dithersubtype == IW_DITHERSUBTYPE_SAMEPATTERN && !is_alpha)
			? ctx->random_seed : ctx->random_seed + s.channeltype;
		d.bkgd1_color_lin = s.bkgd1_color_lin;
		d.bkgd2_color_lin = s.bkgd2_color_lin;
		d.out_cs = (is_alpha || ctx->no_gamma) ? lin : ctx->img2cs;
	}
	p->apply_bkgd = ctx->apply_bkgd;
	p->bkgd_strategy_late = ctx->apply_bkgd_strategy == IWX_BKGD_LATE;
	p->bkgd_checkerboard = ctx->bkgd_checkerboard;
	p->bkgd_check_size = ctx->bkgd_check_size;
	p->bkgd_check_origin[0] = ctx->bkgd_check_origin[0];
	p->bkgd_check_origin[1] = ctx->bkgd_check_origin[1];
	p->bkgd_has_transparency = bkgd_has_transparency(ctx);
	p->bkgd1alpha = ctx->bkgd1alpha;
	p->bkgd2alpha = ctx->bkgd2alpha;
	p->precision = ctx->precision;
}

// ---- entry points --------------------------------------------------------

static int begin_use(iw_context *ctx)
{
	if (ctx->use_count > 0) {
		iw_set_error(ctx, "Internal: Incorrect attempt to reprocess image");
		return 0;
	}
	ctx->use_count++;
	return 1;
}

extern "C" int iwamd_prepare(struct iw_context *ctx, struct iwamd_plan *plan)
{
	if (!begin_use(ctx)) return 0;
	if (!iwx_prepare_processing(ctx, ctx->canvas_width, ctx->canvas_height)) return 0;
	iwx_fill_plan(ctx, plan);
	ctx->prepared = 1;
	return 1;
}

// Completes a prepared context from its output image, with the device's
// optimiser scan when there is one.
int iwx_finish_image(struct iw_context *ctx, const uint8_t *img2, const iwx_opt_scan *scan)
{
	if (!ctx->prepared) {
		iw_set_error(ctx, "Internal: finish without prepare");
		return 0;
	}
	ctx->img2.pixels = (iw_byte *)iw_malloc_large(ctx, ctx->img2.bpr, ctx->img2.height);
	if (!ctx->img2.pixels) return 0;
	memcpy(ctx->img2.pixels, img2, ctx->img2.bpr * ctx->img2.height);
	ctx->prepared = 0;
	return iwx_finish(ctx, scan);
}

extern "C" int iwamd_finish(struct iw_context *ctx, const uint8_t *img2)
{
	return iwx_finish_image(ctx, img2, nullptr);
}

// prepare -> HIP kernels -> finish (src/imagew-main.c:2250-2272).
extern "C" int iw_process_image(struct iw_context *ctx)
{
	if (ctx->use_count > 0) {
		iw_set_error(ctx, "Internal: Incorrect attempt to reprocess image");
		return 0;
	}
	return iwx_gpu_process(ctx);
}
