// Steps after the pixel pipeline: background-colour label, negation and the
// output optimiser.  They read and rewrite the output raster (img2) on the
// host; all are integer work over bytes except the label quantisation.
//
// Semantics: src/imagew-main.c:1238-1366 (label, negate) and
// src/imagew-opt.c:1-1314 (optimiser).  The optimiser's choices decide what
// iw_get_output_image() returns, so they follow the reference exactly: the
// order of the checks, the early exits of the scans, the colour-key search
// (closest unused value to 192), palette order and sort key.
#include "iw_ctx.h"

#include <cstdlib>
#include <cstring>

// ---- background label (src/imagew-main.c:1238-1322) ----------------------

// Nearest output code to a linear sample, no dithering
// (src/imagew-main.c:769-799).
static unsigned int quantize_label_sample(iw_context *ctx, double v, const struct iw_csdescr *cs,
	double maxcc)
{
	if (v < 0.0) v = 0.0;
	if (v > 1.0) v = 1.0;
	double e = iwx_linear_to_x(v, cs) * maxcc;
	if (e > maxcc) e = maxcc;
	if (e < 0.0) e = 0.0;
	double f = floor(e), c = ceil(e);
	unsigned int fi = (unsigned int)f, ci = (unsigned int)c;
	if (fi == ci) return (unsigned int)(0.5 + f);
	auto to_lin = [&](unsigned int k) {
		if (cs->cstype == IW_CSTYPE_LINEAR) return ((double)k) / maxcc;
		return iwx_x_to_linear(((double)k) / maxcc, cs);
	};
	double lf = to_lin(fi), lc = to_lin(ci);
	double dfl = v - lf, dce = lc - v;
	return (unsigned int)(0.5 + ((dce <= dfl) ? c : f));
}

static void process_bkgd_label(iw_context *ctx)
{
	const unsigned int p = ctx->output_profile;
	if (!(p & (IW_PROFILE_PNG_BKGD | IW_PROFILE_RGB8_BKGD | IW_PROFILE_RGB16_BKGD))) return;
	if (ctx->req.suppress_output_bkgd_label) return;
	struct iw_color clr;
	if (ctx->req.output_bkgd_label_valid) clr = ctx->req.output_bkgd_label;
	else if (ctx->img1_bkgd_label_set) clr = ctx->img1_bkgd_label_lin;
	else return;
	if (ctx->to_grayscale) {
		double g = iwx_color_to_grayscale(ctx, clr.c[0], clr.c[1], clr.c[2]);
		clr.c[0] = clr.c[1] = clr.c[2] = g;
	}
	double maxcolor;
	if (p & IW_PROFILE_RGB8_BKGD) maxcolor = 255.0;
	else if (p & IW_PROFILE_RGB16_BKGD) maxcolor = 65535.0;
	else if (ctx->img2.bit_depth == 8) maxcolor = 255.0;
	else if (ctx->img2.bit_depth == 16) maxcolor = 65535.0;
	else return;
	for (int k = 0; k < 3; k++) {
		unsigned int q = quantize_label_sample(ctx, clr.c[k], &ctx->img2cs, maxcolor);
		ctx->img2.bkgdlabel.c[k] = ((double)q) / maxcolor;
	}
	unsigned int a;
	if (clr.c[3] <= 0.0) a = 0;
	else if (clr.c[3] >= 1.0) a = (unsigned int)maxcolor;
	else a = (unsigned int)(0.5 + clr.c[3] * (unsigned int)maxcolor);
	ctx->img2.bkgdlabel.c[3] = ((double)a) / maxcolor;
	ctx->img2.has_bkgdlabel = 1;
}

// ---- negate (src/imagew-main.c:1324-1366) --------------------------------

static void negate_target_image(iw_context *ctx)
{
	iw_image &im = ctx->img2;
	const int nc = ctx->img2_numchannels;
	for (int ch = 0; ch < nc; ch++) {
		const iwx_ci_out &ci = ctx->img2_ci[ch];
		if (ci.channeltype == IW_CHANNELTYPE_ALPHA) continue;
		for (int j = 0; j < im.height; j++) {
			iw_byte *row = im.pixels + (size_t)j * im.bpr;
			for (int i = 0; i < im.width; i++) {
				if (im.sampletype == IW_SAMPLETYPE_FLOATINGPOINT) {
					float s;
					memcpy(&s, row + (size_t)(nc * i + ch) * 4, 4);
					s = 1.0f - s;
					memcpy(row + (size_t)(nc * i + ch) * 4, &s, 4);
				} else if (im.bit_depth == 8) {
					iw_byte &b = row[nc * i + ch];
					b = (iw_byte)(ci.maxcolorcode_int - b);
				} else if (im.bit_depth == 16) {
					iw_byte *q = row + (size_t)(nc * i + ch) * 2;
					unsigned int n = ci.maxcolorcode_int - (q[0] * 256u + q[1]);
					q[0] = (n & 0xff00) >> 8;
					q[1] = n & 0xff;
				}
			}
		}
	}
}

// ---- optimiser (src/imagew-opt.c) ----------------------------------------

namespace {

// Sample k of pixel i of row j at the current depth.
inline unsigned int sample_at(const iwx_opt_ctx &o, int nc, int i, int j, int k)
{
	const iw_byte *p = o.pixelsptr + (size_t)j * o.bpr;
	if (o.bit_depth == 16) {
		const iw_byte *q = p + (size_t)(i * nc + k) * 2;
		return ((unsigned int)q[0] << 8) | q[1];
	}
	return p[i * nc + k];
}

// One pass recording colour / transparency / 16-bit precision, with the
// reference's early stop once every property it can use is known
// (src/imagew-opt.c:18-284).
// The image types and depths scan_pixels examines (else the reductions are
// skipped).
bool scan_types_ok(const iwx_opt_ctx &o)
{
	const int t = o.imgtype, bd = o.bit_depth;
	if (bd != 8 && bd != 16) return false;
	if (t == IW_IMGTYPE_GRAY && bd == 8) return false;
	return t == IW_IMGTYPE_GRAY || t == IW_IMGTYPE_GRAYA || t == IW_IMGTYPE_RGB || t == IW_IMGTYPE_RGBA;
}

int scan_pixels(iwx_opt_ctx &o)
{
	const int t = o.imgtype, bd = o.bit_depth;
	const bool rgb = (t == IW_IMGTYPE_RGB || t == IW_IMGTYPE_RGBA);
	const bool alpha = IW_IMGTYPE_HAS_ALPHA(t);
	if (bd != 8 && bd != 16) return 0;
	if (t == IW_IMGTYPE_GRAY && bd == 8) return 0;
	if (t != IW_IMGTYPE_GRAY && t != IW_IMGTYPE_GRAYA && !rgb) return 0;
	const int nc = iw_imgtype_num_channels(t);
	const unsigned int amax = (bd == 16) ? 65535u : 255u;
	const bool want16 = (bd == 16);
	for (int j = 0; j < o.height; j++) {
		const iw_byte *row = o.pixelsptr + (size_t)j * o.bpr;
		for (int i = 0; i < o.width; i++) {
			if (want16) {
				const iw_byte *q = row + (size_t)i * nc * 2;
				for (int k = 0; k < nc; k++)
					if (q[2 * k] != q[2 * k + 1]) o.has_16bit_precision = 1;
			}
			if (alpha) {
				unsigned int a = sample_at(o, nc, i, j, nc - 1);
				if (a < amax) {
					o.has_transparency = 1;
					if (a > 0) o.has_partial_transparency = 1;
				}
			}
			if (rgb) {
				unsigned int r = sample_at(o, nc, i, j, 0), g = sample_at(o, nc, i, j, 1),
					b = sample_at(o, nc, i, j, 2);
				if (r != g || r != b) o.has_color = 1;
			}
			// Stop once nothing more can change the outcome.
			bool done;
			if (t == IW_IMGTYPE_RGB) done = want16 ? (o.has_color && o.has_16bit_precision) : o.has_color;
			else if (t == IW_IMGTYPE_RGBA)
				done = want16 ? (o.has_partial_transparency && o.has_color && o.has_16bit_precision)
					: (o.has_partial_transparency && o.has_color);
			else if (t == IW_IMGTYPE_GRAYA)
				done = want16 ? (o.has_partial_transparency && o.has_16bit_precision)
					: o.has_partial_transparency;
			else done = o.has_16bit_precision;
			if (done) return 1;
		}
	}
	return 1;
}

void attach_pixels(iw_context *ctx, iwx_opt_ctx &o, iw_byte *np, size_t nbpr)
{
	if (o.tmp_pixels) iw_free(ctx, o.tmp_pixels);
	o.tmp_pixels = np;
	o.pixelsptr = np;
	o.bpr = nbpr;
}

// Keep the high byte of every sample (src/imagew-opt.c:243-283).
void reduce_16_to_8(iw_context *ctx, iwx_opt_ctx &o, int spp)
{
	if (!ctx->opt_16_to_8) return;
	size_t nbpr = iw_calc_bytesperrow(o.width, 8 * spp);
	iw_byte *np = (iw_byte *)iw_malloc_large(ctx, nbpr, o.height);
	if (!np) return;
	for (int j = 0; j < o.height; j++)
		for (int i = 0; i < spp * o.width; i++)
			np[(size_t)j * nbpr + i] = o.pixelsptr[(size_t)j * o.bpr + (size_t)i * 2];
	attach_pixels(ctx, o, np, nbpr);
	o.bit_depth = 8;
	if (o.has_bkgdlabel)
		for (int k = 0; k < 4; k++) o.bkgdlabel[k] >>= 8;
}

// New image from up to three channels of the current one
// (src/imagew-opt.c:286-360).
void copy_channels(iw_context *ctx, iwx_opt_ctx &o, int new_type, int c0, int c1, int c2)
{
	const int oldnc = iw_imgtype_num_channels(o.imgtype), newnc = iw_imgtype_num_channels(new_type);
	const int bps = o.bit_depth / 8;
	size_t nbpr = iw_calc_bytesperrow(o.width, o.bit_depth * newnc);
	iw_byte *np = (iw_byte *)iw_malloc_large(ctx, nbpr, o.height);
	if (!np) return;
	const int src[3] = {c0, c1, c2};
	for (int j = 0; j < o.height; j++)
		for (int i = 0; i < o.width; i++)
			for (int k = 0; k < newnc; k++)
				for (int b = 0; b < bps; b++)
					np[(size_t)j * nbpr + (size_t)(i * newnc + k) * bps + b] =
						o.pixelsptr[(size_t)j * o.bpr + (size_t)(i * oldnc + src[k]) * bps + b];
	attach_pixels(ctx, o, np, nbpr);
	o.imgtype = new_type;
}

// The unused value in flags[0..count) nearest to 192 (src/imagew-opt.c:389-414).
int find_unused(const iw_byte *flags, int count, iw_byte *out)
{
	int found = 0, best = 0;
	for (int i = 0; i < count; i++) {
		if (flags[i]) continue;
		int d = abs(i - 192);
		if (!found || d < best) {
			*out = (iw_byte)i;
			found = 1;
			best = d;
		}
	}
	if (!found) *out = 0;
	return found;
}

// RGBA/GA with only fully opaque or fully transparent pixels becomes RGB/G
// plus a colour key chosen among values the image does not use
// (src/imagew-opt.c:418-683).
void try_binary_trns(iw_context *ctx, iwx_opt_ctx &o, const iwx_opt_scan *scan)
{
	if (!(ctx->output_profile & IW_PROFILE_BINARYTRNS)) return;
	if (!ctx->opt_binary_trns) return;
	const bool rgb = (o.imgtype == IW_IMGTYPE_RGBA);
	const bool b16 = (o.bit_depth == 16);
	const int nc = rgb ? 4 : 2;
	iw_byte used[256];
	memset(used, 0, sizeof(used));
	const bool have_keys = scan && (scan->valid & IWX_SCAN_HAVE_KEYS);
	if (have_keys) {
		// The device's sets of key components in use (the same tests below).
		const int key = rgb ? (b16 ? IWX_KEY_RGB16 : IWX_KEY_RGB8) : (b16 ? IWX_KEY_GRAY16 : IWX_KEY_GRAY8);
		for (int v = 0; v < 256; v++) used[v] = (scan->used[key][v >> 5] >> (v & 31)) & 1u;
	}
	iw_byte *mask = (iw_byte *)iw_malloc_large(ctx, o.width, o.height);
	if (!mask) return;
	for (int j = 0; j < o.height; j++) {
		for (int i = 0; i < o.width; i++) {
			const unsigned int a = sample_at(o, nc, i, j, nc - 1);
			mask[(size_t)j * o.width + i] = (a != 0);
			if (a == 0 || have_keys) continue;
			if (rgb) {
				const unsigned int r = sample_at(o, nc, i, j, 0), g = sample_at(o, nc, i, j, 1),
					b = sample_at(o, nc, i, j, 2);
				if (b16) {
					if ((r >> 8) != 192 || g != 192 * 257 || b != 192 * 257) continue;
					used[r & 0xff] = 1;
				} else {
					if (g != 192 || b != 192) continue;
					used[r] = 1;
				}
			} else {
				const unsigned int v = sample_at(o, nc, i, j, 0);
				if (b16) {
					if ((v >> 8) != 192) continue;
					used[v & 0xff] = 1;
				} else {
					used[v] = 1;
				}
			}
		}
	}
	iw_byte key = 0;
	if (find_unused(used, 256, &key)) {
		copy_channels(ctx, o, rgb ? IW_IMGTYPE_RGB : IW_IMGTYPE_GRAY, 0, rgb ? 1 : 0, rgb ? 2 : 0);
		if (o.tmp_pixels) {
			const int nnc = rgb ? 3 : 1;
			for (int j = 0; j < o.height; j++) {
				for (int i = 0; i < o.width; i++) {
					if (mask[(size_t)j * o.width + i]) continue;
					iw_byte *q = o.tmp_pixels + (size_t)j * o.bpr + (size_t)i * nnc * (b16 ? 2 : 1);
					if (b16) {
						for (int k = 0; k < 2 * nnc; k++) q[k] = 192;
						q[1] = key;
					} else {
						q[0] = key;
						if (rgb) { q[1] = 192; q[2] = 192; }
					}
				}
			}
			o.has_colorkey_trns = 1;
			unsigned int kr = b16 ? 192 * 256 + key : key;
			unsigned int kgb = rgb ? (b16 ? 192 * 256 + 192 : 192) : kr;
			o.colorkey[IW_CHANNELTYPE_RED] = kr;
			o.colorkey[IW_CHANNELTYPE_GREEN] = kgb;
			o.colorkey[IW_CHANNELTYPE_BLUE] = kgb;
		}
	}
	iw_free(ctx, mask);
}

// ---- palettes (src/imagew-opt.c:678-1094) --------------------------------

struct iw_rgba8color pixel_color(const iwx_opt_ctx &o, int x, int y)
{
	const int spp = iw_imgtype_num_channels(o.imgtype);
	const iw_byte *p = o.pixelsptr + (size_t)y * o.bpr + (size_t)x * spp;
	struct iw_rgba8color c;
	switch (o.imgtype) {
	case IW_IMGTYPE_RGB: c = {p[0], p[1], p[2], 255}; break;
	case IW_IMGTYPE_RGBA: c = {p[0], p[1], p[2], p[3]}; break;
	case IW_IMGTYPE_GRAYA: c = {p[0], p[0], p[0], p[1]}; break;
	default: c = {p[0], p[0], p[0], 255}; break;
	}
	if ((o.imgtype == IW_IMGTYPE_RGBA || o.imgtype == IW_IMGTYPE_GRAYA) && c.a == 0) c.r = c.g = c.b = 0;
	return c;
}

int palette_find(const struct iw_palette *pal, const struct iw_rgba8color &c)
{
	for (int i = 0; i < pal->num_entries; i++) {
		const auto &e = pal->entry[i];
		if (e.r == c.r && e.g == c.g && e.b == c.b && e.a == c.a) return i;
	}
	return -1;
}

// The distinct colours of the image, in the order of first appearance, or 0
// when there are more than 256.  With the device's set the order differs;
// every later use sorts the entries or only tests membership
// (src/imagew-opt.c:871-899, :906-1004), so the outcome is the same.
int collect_palette(iwx_opt_ctx &o, const iwx_opt_scan *scan)
{
	struct iw_palette *pal = o.palette;
	if (scan && (scan->valid & IWX_SCAN_HAVE_PAL)) {
		if (scan->pal_over) return 0;
		for (int s = 0; s < IWX_SCAN_PAL_SLOTS; s++) {
			const unsigned long long v = scan->pal[s];
			if (!v) continue;
			if (pal->num_entries >= 256) return 0;
			pal->entry[pal->num_entries++] = {(iw_byte)v, (iw_byte)(v >> 8), (iw_byte)(v >> 16), (iw_byte)(v >> 24)};
		}
	} else {
		for (int y = 0; y < o.height; y++) {
			for (int x = 0; x < o.width; x++) {
				struct iw_rgba8color c = pixel_color(o, x, y);
				if (palette_find(pal, c) >= 0) continue;
				if (pal->num_entries >= 256) return 0;
				pal->entry[pal->num_entries++] = c;
			}
		}
	}
	if (pal->num_entries < 1) return 0;
	if (o.has_bkgdlabel) {
		struct iw_rgba8color c = {(iw_byte)o.bkgdlabel[0], (iw_byte)o.bkgdlabel[1],
			(iw_byte)o.bkgdlabel[2], 255};
		// A fully transparent entry, or one with the label's RGB, serves.
		int found = -1;
		for (int i = 0; i < pal->num_entries && found < 0; i++) {
			const auto &e = pal->entry[i];
			if (e.a == 0 || (e.r == c.r && e.g == c.g && e.b == c.b)) found = i;
		}
		if (found < 0) {
			if (pal->num_entries >= 256) return 0;
			pal->entry[pal->num_entries++] = c;
		}
	}
	return 1;
}

int palette_opt_ok(iw_context *ctx, iwx_opt_ctx &o, int bpp)
{
	if (!ctx->opt_palette) return 0;
	if (o.has_transparency && !(ctx->output_profile & IW_PROFILE_PALETTETRNS)) return 0;
	unsigned int flag;
	int maxn;
	switch (bpp) {
	case 1: flag = IW_PROFILE_PAL1; maxn = 2; break;
	case 2: flag = IW_PROFILE_PAL2; maxn = 4; break;
	case 4: flag = IW_PROFILE_PAL4; maxn = 16; break;
	case 8: flag = IW_PROFILE_PAL8; maxn = 256; break;
	default: return 0;
	}
	return (ctx->output_profile & flag) && o.palette->num_entries <= maxn;
}

int palette_is_valid_gray(iw_context *ctx, iwx_opt_ctx &o, int bpp, int *binary_trns,
	unsigned int *trns_shade)
{
	*binary_trns = 0;
	*trns_shade = 0;
	if (o.has_color || o.has_partial_transparency) return 0;
	if (o.has_transparency && !ctx->opt_binary_trns) return 0;
	if (o.has_transparency && !(ctx->output_profile & IW_PROFILE_BINARYTRNS)) return 0;
	int factor, maxn;
	switch (bpp) {
	case 1: factor = 255; maxn = 2; break;
	case 2: factor = 85; maxn = 4; break;
	case 4: factor = 17; maxn = 16; break;
	case 8: factor = 1; maxn = 256; break;
	default: return 0;
	}
	if (o.palette->num_entries > maxn) return 0;
	if (o.has_bkgdlabel && bpp < 8 && o.bkgdlabel[0] % factor != 0) return 0;
	iw_byte used[256];
	memset(used, 0, sizeof(used));
	for (int i = 0; i < o.palette->num_entries; i++) {
		const auto &e = o.palette->entry[i];
		if (e.a == 0) continue;
		if (e.r % factor) return 0;
		used[e.r / factor] = 1;
	}
	if (o.has_transparency) {
		iw_byte key = 0;
		if (!find_unused(used, maxn, &key)) return 0;
		*binary_trns = 1;
		*trns_shade = key;
	}
	return 1;
}

void make_gray_palette(iwx_opt_ctx &o, int bpp)
{
	int factor;
	switch (bpp) {
	case 1: factor = 255; o.palette->num_entries = 2; break;
	case 2: factor = 85; o.palette->num_entries = 4; break;
	case 4: factor = 17; o.palette->num_entries = 16; break;
	default: factor = 1; o.palette->num_entries = 256; break;
	}
	for (int i = 0; i < o.palette->num_entries; i++) {
		iw_byte v = (iw_byte)(factor * i);
		o.palette->entry[i] = {v, v, v, 255};
	}
	o.palette_is_grayscale = 1;
}

// Translucent entries first, then by r+g+b, then g, r, b, a
// (src/imagew-opt.c:871-899).
int palette_order(const void *p1, const void *p2)
{
	const auto *a = (const struct iw_rgba8color *)p1, *b = (const struct iw_rgba8color *)p2;
	if (a->a == 255 && b->a < 255) return 1;
	if (a->a < 255 && b->a == 255) return -1;
	int s1 = a->r + a->g + a->b, s2 = b->r + b->g + b->b;
	if (s1 != s2) return s1 > s2 ? 1 : -1;
	if (a->g != b->g) return a->g > b->g ? 1 : -1;
	if (a->r != b->r) return a->r > b->r ? 1 : -1;
	if (a->b != b->b) return a->b > b->b ? 1 : -1;
	if (a->a != b->a) return a->a > b->a ? 1 : -1;
	return 0;
}

// Colour -> palette index, the first entry that matches as palette_find
// gives it, through an open-addressed table of the at most 256 entries.
struct PaletteIndex {
	uint32_t key[1024];
	int16_t idx[1024];
	static uint32_t pack(const struct iw_rgba8color &c)
	{
		return (uint32_t)c.r | ((uint32_t)c.g << 8) | ((uint32_t)c.b << 16) | ((uint32_t)c.a << 24);
	}
	static unsigned slot(uint32_t k) { return (k * 2654435761u) >> 22; }
	explicit PaletteIndex(const struct iw_palette *pal)
	{
		memset(idx, 0xff, sizeof(idx));
		for (int i = 0; i < pal->num_entries; i++) {
			const uint32_t k = pack(pal->entry[i]);
			unsigned s = slot(k);
			while (idx[s] >= 0 && key[s] != k) s = (s + 1) & 1023;
			if (idx[s] < 0) {
				key[s] = k;
				idx[s] = (int16_t)i;
			}
		}
	}
	int find(const struct iw_rgba8color &c) const
	{
		const uint32_t k = pack(c);
		for (unsigned s = slot(k);; s = (s + 1) & 1023) {
			if (idx[s] < 0) return -1;
			if (key[s] == k) return idx[s];
		}
	}
};

void convert_to_palette(iw_context *ctx, iwx_opt_ctx &o)
{
	size_t nbpr = o.width;
	iw_byte *np = (iw_byte *)iw_malloc_large(ctx, nbpr, o.height);
	if (!np) return;
	const PaletteIndex index(o.palette);
	for (int y = 0; y < o.height; y++) {
		for (int x = 0; x < o.width; x++) {
			struct iw_rgba8color c = pixel_color(o, x, y);
			int e;
			if (o.has_colorkey_trns && c.a == 0) e = (int)o.colorkey[IW_CHANNELTYPE_RED];
			else {
				e = index.find(c);
				if (e < 0) e = 0;
			}
			np[(size_t)y * nbpr + x] = (iw_byte)e;
		}
	}
	attach_pixels(ctx, o, np, nbpr);
	o.bit_depth = 8;
	o.imgtype = IW_IMGTYPE_PALETTE;
}

// Preference gray1, pal1, gray2, pal2, gray4, pal4, (gray8: leave), pal8
// (src/imagew-opt.c:1010-1094).
void try_palette(iw_context *ctx, iwx_opt_ctx &o, const iwx_opt_scan *scan)
{
	const unsigned int p = ctx->output_profile;
	if (!(p & (IW_PROFILE_PAL1 | IW_PROFILE_PAL2 | IW_PROFILE_PAL4 | IW_PROFILE_PAL8 |
			   IW_PROFILE_GRAY1 | IW_PROFILE_GRAY2 | IW_PROFILE_GRAY4)))
		return;
	if (o.bit_depth != 8) return;
	o.palette = (struct iw_palette *)iw_malloc(ctx, sizeof(struct iw_palette));
	if (!o.palette) return;
	o.palette->num_entries = 0;
	int bt = 0;
	unsigned int shade = 0;
	bool chosen = false;
	if (collect_palette(o, scan)) {
		static const int bpps[3] = {1, 2, 4};
		static const unsigned int grayflag[3] = {IW_PROFILE_GRAY1, IW_PROFILE_GRAY2, IW_PROFILE_GRAY4};
		bool stop = false;
		for (int k = 0; k < 3 && !chosen && !stop; k++) {
			if (ctx->opt_grayscale && (p & grayflag[k]) && palette_is_valid_gray(ctx, o, bpps[k], &bt, &shade)) {
				make_gray_palette(o, bpps[k]);
				if (bt) {
					o.has_colorkey_trns = 1;
					o.colorkey[0] = o.colorkey[1] = o.colorkey[2] = shade;
				}
				chosen = true;
			} else if (palette_opt_ok(ctx, o, bpps[k])) {
				chosen = true;
			}
		}
		if (!chosen) {
			if (ctx->opt_grayscale && (p & IW_PROFILE_GRAYSCALE) && palette_is_valid_gray(ctx, o, 8, &bt, &shade)) {
				// Best stored as 8-bit grayscale; not this optimisation's job.
			} else if (palette_opt_ok(ctx, o, 8)) {
				chosen = true;
			}
		}
		if (chosen) {
			if (!o.palette_is_grayscale)
				qsort(o.palette->entry, o.palette->num_entries, sizeof(struct iw_rgba8color), palette_order);
			convert_to_palette(ctx, o);
		}
	}
	if (o.imgtype != IW_IMGTYPE_PALETTE) {
		iw_free(ctx, o.palette);
		o.palette = nullptr;
	}
}

// Zero the colour of fully transparent pixels (src/imagew-opt.c:1112-1163).
void make_transparent_pixels_black(iw_image &im)
{
	if (!IW_IMGTYPE_HAS_ALPHA(im.imgtype)) return;
	const int nc = iw_imgtype_num_channels(im.imgtype);
	const int bps = im.bit_depth > 8 ? 2 : 1;
	for (int j = 0; j < im.height; j++) {
		for (int i = 0; i < im.width; i++) {
			iw_byte *p = im.pixels + (size_t)j * im.bpr + (size_t)i * nc * bps;
			bool transparent = (bps == 2) ? (p[(nc - 1) * 2] == 0 && p[(nc - 1) * 2 + 1] == 0) : p[nc - 1] == 0;
			if (transparent) memset(p, 0, (size_t)(nc - 1) * bps);
		}
	}
}

}  // namespace

void iwx_optimize_image(iw_context *ctx, const iwx_opt_scan *scan)
{
	if (scan && !(scan->valid & IWX_SCAN_HAVE_FLAGS)) scan = nullptr;
	iwx_opt_ctx &o = ctx->optctx;
	o.width = ctx->img2.width;
	o.height = ctx->img2.height;
	o.imgtype = ctx->img2.imgtype;
	o.bit_depth = ctx->img2.bit_depth;
	o.bpr = ctx->img2.bpr;
	o.pixelsptr = ctx->img2.pixels;
	if (ctx->img2.has_bkgdlabel) {
		o.has_bkgdlabel = 1;
		for (int k = 0; k < 4; k++)
			o.bkgdlabel[k] = iw_color_get_int_sample(&ctx->img2.bkgdlabel, k, ctx->img2.bit_depth == 8 ? 255 : 65535);
	}
	if (ctx->img2.sampletype != IW_SAMPLETYPE_UINT) return;
	if (ctx->reduced_output_maxcolor_flag) return;

	if (!scan) make_transparent_pixels_black(ctx->img2);

	if (o.has_bkgdlabel) {
		if (o.bkgdlabel[0] != o.bkgdlabel[1] || o.bkgdlabel[0] != o.bkgdlabel[2]) o.has_color = 1;
		if (o.bit_depth == 16 && (o.bkgdlabel[0] % 257 || o.bkgdlabel[1] % 257 || o.bkgdlabel[2] % 257))
			o.has_16bit_precision = 1;
	}

	if (scan ? scan_types_ok(o) : scan_pixels(o)) {
		if (scan) {
			if (scan->flags & IWX_SCAN_COLOR) o.has_color = 1;
			if (scan->flags & IWX_SCAN_TRNS) o.has_transparency = 1;
			if (scan->flags & IWX_SCAN_PARTIAL) o.has_partial_transparency = 1;
			if (scan->flags & IWX_SCAN_16BIT) o.has_16bit_precision = 1;
		}
		const int t0 = o.imgtype;
		if (o.bit_depth == 16 && !o.has_16bit_precision) reduce_16_to_8(ctx, o, iw_imgtype_num_channels(t0));
		if (!o.has_transparency && ctx->opt_strip_alpha) {
			if (o.imgtype == IW_IMGTYPE_RGBA) copy_channels(ctx, o, IW_IMGTYPE_RGB, 0, 1, 2);
			else if (o.imgtype == IW_IMGTYPE_GRAYA) copy_channels(ctx, o, IW_IMGTYPE_GRAY, 0, 0, 0);
		}
		if (!o.has_color && (ctx->output_profile & IW_PROFILE_GRAYSCALE) && ctx->opt_grayscale) {
			if (o.imgtype == IW_IMGTYPE_RGB) copy_channels(ctx, o, IW_IMGTYPE_GRAY, 0, 0, 0);
			else if (o.imgtype == IW_IMGTYPE_RGBA) copy_channels(ctx, o, IW_IMGTYPE_GRAYA, 0, 3, 0);
		}
	}

	try_palette(ctx, o, scan);

	if ((o.imgtype == IW_IMGTYPE_RGBA || o.imgtype == IW_IMGTYPE_GRAYA) && !o.has_partial_transparency &&
		(o.bit_depth == 8 || o.bit_depth == 16))
		try_binary_trns(ctx, o, scan);
}

int iwx_finish(iw_context *ctx, const iwx_opt_scan *scan)
{
	process_bkgd_label(ctx);
	// The device scans only images that are not negated afterwards.
	if (ctx->req.negate_target) {
		negate_target_image(ctx);
		scan = nullptr;
	}
	iwx_optimize_image(ctx, scan);
	return 1;
}
