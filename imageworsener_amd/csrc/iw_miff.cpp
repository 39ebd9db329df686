// MIFF reading and writing for floating-point images: the format the
// reference uses to carry an image between passes without clamping or
// quantising it (HDRI profile).
//
// Behaviour follows src/imagew-miff.c (:1-897): a text header of
// name=value attributes (values may be {braced}) ended by ':' and one more
// byte, then rows of big-endian samples.  Reading accepts 32- and 64-bit
// floating-point DirectClass images, RGB or Gray, with or without alpha
// ("matte"), uncompressed (Zip needs zlib, which this library does not carry,
// so it is refused as the reference refuses it when built without zlib);
// 64-bit samples become 32-bit floats.  Writing produces the reference's
// header text and uncompressed 32-bit rows, so the bytes of a written file
// are the reference's.
#include "iw_ctx.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct MiffReader {
	iw_context *ctx;
	iw_iodescr *io;
	bool failed = false;  // a read came up short
	bool bad = false;     // an error was reported

	bool read(void *buf, size_t n)
	{
		size_t got = 0;
		if (!io->read_fn(ctx, io, buf, n, &got) || got != n) {
			failed = true;
			return false;
		}
		return true;
	}
	int byte()
	{
		iw_byte b = 0;
		return read(&b, 1) ? (int)b : -1;
	}
	void error(const char *msg)
	{
		if (!bad) iw_set_error(ctx, msg);
		bad = true;
	}
};

struct MiffHeader {
	int width = 0, height = 0, depth = 0;
	bool alpha = false, gray = false;
	bool density_known = false, per_cm = false;
	double dx = 0.0, dy = 0.0;
	int intent = 0;
	long long profile_bytes = 0;
	struct iw_csdescr cs;
};

double hex4(const char *s)
{
	int v = 0;
	for (int k = 0; k < 4; k++) {
		const char d = s[k];
		const int h = (d >= '0' && d <= '9') ? d - '0' : (d >= 'a' && d <= 'f') ? d - 'a' + 10 :
			(d >= 'A' && d <= 'F') ? d - 'A' + 10 : 0;
		v = v * 16 + h;
	}
	return (double)v / 65535.0;
}

// One attribute of the header (src/imagew-miff.c:193-302).
void miff_attribute(MiffReader &r, MiffHeader &h, const std::string &name, const std::string &val)
{
	if (r.bad) return;
	const char *v = val.c_str();
	if (name == "matte") {
		if (v[0] == 'T') h.alpha = true;
	} else if (name == "class") {
		if (iw_stricmp(v, "DirectClass")) r.error("MIFF: Unsupported image class");
	} else if (name == "columns") {
		h.width = atoi(v);
	} else if (name == "rows") {
		h.height = atoi(v);
	} else if (name == "colorspace") {
		if (!iw_stricmp(v, "RGB")) {
		} else if (!iw_stricmp(v, "sRGB")) {
			iw_make_srgb_csdescr_2(&h.cs);
		} else if (!iw_stricmp(v, "Gray")) {
			h.gray = true;
		} else {
			r.error("MIFF: Unsupported colorspace");
		}
	} else if (name == "rendering-intent") {
		if (!iw_stricmp(v, "perceptual")) h.intent = IW_INTENT_PERCEPTUAL;
		else if (!iw_stricmp(v, "relative")) h.intent = IW_INTENT_RELATIVE;
		else if (!iw_stricmp(v, "saturation")) h.intent = IW_INTENT_SATURATION;
		else if (!iw_stricmp(v, "absolute")) h.intent = IW_INTENT_ABSOLUTE;
		else iw_warning(r.ctx, "MIFF: Unrecognized rendering-intent");
	} else if (name == "depth") {
		h.depth = atoi(v);
		if (h.depth != 32 && h.depth != 64) {
			if (!r.bad) iw_set_errorf(r.ctx, "MIFF: Unsupported bit depth (%d)", h.depth);
			r.bad = true;
		}
	} else if (name == "compression") {
		if (!iw_stricmp(v, "None")) {
		} else if (!iw_stricmp(v, "Zip")) {
			r.error("MIFF: Zip compression is not supported");
		} else {
			r.error("MIFF: Unsupported compression");
		}
	} else if (name == "gamma") {
		const double g = atof(v);
		if (g >= 0.00001 && g <= 10.0) iw_make_gamma_csdescr(&h.cs, 1.0 / g);
	} else if (name == "background-color") {
		// Only #rrrrggggbbbb[aaaa].
		if ((val.size() == 13 || val.size() == 17) && v[0] == '#') {
			struct iw_color c;
			c.c[0] = hex4(v + 1);
			c.c[1] = hex4(v + 5);
			c.c[2] = hex4(v + 9);
			c.c[3] = val.size() == 17 ? hex4(v + 13) : 1.0;
			iw_set_input_bkgd_label_2(r.ctx, &c);
		}
	} else if (name == "quantum:format") {
		if (iw_stricmp(v, "floating-point")) r.error("MIFF: Unsupported sample format");
	} else if (name == "units") {
		if (!iw_stricmp(v, "PixelsPerCentimeter")) h.per_cm = true;
	} else if (name == "resolution") {
		const char *x = strchr(v, 'x');
		if (x) {
			h.dx = atof(v);
			h.dy = atof(x + 1);
			h.density_known = true;
		}
	} else if (name.compare(0, 8, "profile:") == 0) {
		// Profile data follows the header; it is skipped, so its size counts.
		h.profile_bytes += atoi(v);
	}
}

// The header (src/imagew-miff.c:304-395): names and values up to 100
// characters, a value in braces may hold spaces, ':' where a name would start
// ends the header, and the byte after it is skipped.
bool miff_header(MiffReader &r, MiffHeader &h)
{
	enum { NEUTRAL, NAME, VALUE, QUOTED } st = NEUTRAL;
	std::string name, val;
	for (;;) {
		if (r.bad) return false;
		const int c = r.byte();
		if (c < 0) return false;
		const char b = (char)c;
		const bool ws = b >= 0 && b <= 32;
		if (st == NEUTRAL) {
			if (ws) continue;
			st = NAME;
		}
		if (st == NAME) {
			if (name.empty() && b == ':') break;
			if (b == '=') {
				st = VALUE;
				continue;
			}
			if (name.size() < 100) name.push_back(b);
		}
		if (st == VALUE && val.empty() && b == '{') {
			st = QUOTED;
			continue;
		}
		if ((st == VALUE && ws) || (st == QUOTED && b == '}')) {
			miff_attribute(r, h, name, val);
			name.clear();
			val.clear();
			st = NEUTRAL;
			continue;
		}
		if (st == VALUE || st == QUOTED) {
			if (val.size() < 100) val.push_back(b);
		}
	}
	(void)r.byte();
	return !r.failed && !r.bad;
}

float be_float32(const iw_byte *p)
{
	const uint32_t u = ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
	float f;
	memcpy(&f, &u, 4);
	return f;
}

double be_float64(const iw_byte *p)
{
	uint64_t u = 0;
	for (int k = 0; k < 8; k++) u = (u << 8) | p[k];
	double d;
	memcpy(&d, &u, 8);
	return d;
}

}  // namespace

extern "C" int iw_read_miff_file(struct iw_context *ctx, struct iw_iodescr *io)
{
	MiffReader r{ctx, io};
	MiffHeader h;
	iw_make_srgb_csdescr_2(&h.cs);  // sRGB unless the header says otherwise
	struct iw_image img;
	memset(&img, 0, sizeof(img));
	bool ok = miff_header(r, h);
	if (ok && h.depth != 32 && h.depth != 64) {
		iw_set_error(ctx, "MIFF: Unsupported or unset bit depth");
		ok = false;
	}
	ok = ok && iw_check_image_dimensions(ctx, h.width, h.height);
	if (ok) {
		const int nch = (h.gray ? 1 : 3) + (h.alpha ? 1 : 0);
		img.imgtype = h.gray ? (h.alpha ? IW_IMGTYPE_GRAYA : IW_IMGTYPE_GRAY) :
			(h.alpha ? IW_IMGTYPE_RGBA : IW_IMGTYPE_RGB);
		img.native_grayscale = h.gray ? 1 : 0;
		img.sampletype = IW_SAMPLETYPE_FLOATINGPOINT;
		img.bit_depth = 32;
		img.width = h.width;
		img.height = h.height;
		img.rendering_intent = h.intent;
		if (h.density_known && h.per_cm) {
			img.density_x = h.dx * 100.0;
			img.density_y = h.dy * 100.0;
			img.density_code = IW_DENSITY_UNITS_PER_METER;
		}
		// Skip the profile data.
		std::vector<iw_byte> skip(2048);
		for (long long left = h.profile_bytes; ok && left > 0;) {
			const size_t n = (size_t)(left < 2048 ? left : 2048);
			ok = r.read(skip.data(), n);
			left -= (long long)n;
		}
		const size_t nsamp = (size_t)nch * (size_t)h.width;
		const size_t fbytes = (size_t)(h.depth / 8);
		std::vector<iw_byte> row(nsamp * fbytes);
		img.bpr = nsamp * 4;
		img.pixels = ok ? (iw_byte *)iw_malloc_large(ctx, img.bpr, img.height) : nullptr;
		ok = ok && img.pixels;
		for (int j = 0; ok && j < h.height; j++) {
			ok = r.read(row.data(), row.size());
			if (!ok) break;
			iw_byte *dst = img.pixels + (size_t)j * img.bpr;
			for (size_t i = 0; i < nsamp; i++) {
				const float f = h.depth == 32 ? be_float32(&row[i * 4]) : (float)be_float64(&row[i * 8]);
				memcpy(dst + i * 4, &f, 4);
			}
		}
	}
	if (!ok) {
		iw_set_error(ctx, "Failed to read MIFF file");
		iw_free(ctx, img.pixels);
		return 0;
	}
	iw_set_input_image(ctx, &img);
	iw_set_input_colorspace(ctx, &h.cs);
	return 1;
}

extern "C" int iw_write_miff_file(struct iw_context *ctx, struct iw_iodescr *io)
{
	struct iw_image img;
	iw_get_output_image(ctx, &img);
	if (img.sampletype != IW_SAMPLETYPE_FLOATINGPOINT) {
		iw_set_error(ctx, "Internal: Bad image type for MIFF");
		return 0;
	}
	int nch;
	bool alpha = false;
	switch (img.imgtype) {
	case IW_IMGTYPE_GRAY: nch = 1; break;
	case IW_IMGTYPE_GRAYA: nch = 2; alpha = true; break;
	case IW_IMGTYPE_RGB: nch = 3; break;
	case IW_IMGTYPE_RGBA: nch = 4; alpha = true; break;
	default: return 0;
	}
	// Without zlib every row is stored uncompressed, whatever was asked for.
	std::string hdr = "id=ImageMagick  version=1.0\n";
	char buf[200];
	snprintf(buf, sizeof buf, "class=DirectClass  colors=0  matte=%s\n", alpha ? "True" : "False");
	hdr += buf;
	snprintf(buf, sizeof buf, "columns=%d  rows=%d  depth=%d\n", img.width, img.height, img.bit_depth);
	hdr += buf;
	if (img.imgtype == IW_IMGTYPE_GRAY) hdr += "type=Grayscale\ncolorspace=Gray\n";
	else if (img.imgtype == IW_IMGTYPE_GRAYA) hdr += "type=GrayscaleMatte\ncolorspace=Gray\n";
	else hdr += "colorspace=RGB\n";
	hdr += "compression=None  quality=0\n";
	if (img.density_code == IW_DENSITY_UNITS_PER_METER) {
		hdr += "units=PixelsPerCentimeter\n";
		iw_snprintf(buf, sizeof buf, "resolution=%.2fx%.2f\n", img.density_x / 100.0, img.density_y / 100.0);
		hdr += buf;
	}
	const char *intent = img.rendering_intent == IW_INTENT_PERCEPTUAL ? "Perceptual" :
		img.rendering_intent == IW_INTENT_RELATIVE ? "Relative" :
		img.rendering_intent == IW_INTENT_SATURATION ? "Saturation" :
		img.rendering_intent == IW_INTENT_ABSOLUTE ? "Absolute" : nullptr;
	if (intent) {
		snprintf(buf, sizeof buf, "rendering-intent=%s\n", intent);
		hdr += buf;
	}
	hdr += "gamma=1.0\n";
	if (img.has_bkgdlabel) {
		unsigned int k[4];
		for (int c = 0; c < 4; c++) k[c] = iw_color_get_int_sample(&img.bkgdlabel, c, 65535);
		if (k[3] >= 65535) snprintf(buf, sizeof buf, "background-color=#%04x%04x%04x\n", k[0], k[1], k[2]);
		else snprintf(buf, sizeof buf, "background-color=#%04x%04x%04x%04x\n", k[0], k[1], k[2], k[3]);
		hdr += buf;
	}
	hdr += "quantum:format={floating-point}\n";
	hdr.append("\x0c\x0a\x3a\x1a", 4);
	io->write_fn(ctx, io, hdr.data(), hdr.size());
	// Rows of big-endian 32-bit samples.
	const size_t nsamp = (size_t)nch * (size_t)img.width;
	const size_t bps = (size_t)(img.bit_depth / 8);
	std::vector<iw_byte> row(bps * nsamp, 0);
	for (int j = 0; j < img.height; j++) {
		const iw_byte *src = img.pixels + (size_t)j * img.bpr;
		for (size_t i = 0; i < nsamp && bps == 4; i++) {
			uint32_t u;
			memcpy(&u, src + i * 4, 4);
			for (int b = 0; b < 4; b++) row[i * 4 + b] = (iw_byte)(u >> (24 - 8 * b));
		}
		io->write_fn(ctx, io, row.data(), row.size());
	}
	return 1;
}
