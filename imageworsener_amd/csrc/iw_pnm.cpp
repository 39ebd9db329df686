// PNM (PBM/PGM/PPM, binary) and PAM reading and writing, and the format
// dispatch iw_read_file_by_fmt / iw_write_file_by_fmt (PNM, PAM, and MIFF
// through iw_miff.cpp).
//
// The dependency-free on-disk format on either side of the pixel path.
// Behaviour follows the reference's src/imagew-pnm.c (:1-958) and
// src/imagew-allfmts.c (:18-182): binary bitmaps only; a MAXVAL other than
// 255/65535 becomes a reduced input max colour code; 16-bit samples are
// big-endian like the library's internal layout, so rows copy straight in and
// out; PBM bits are inverted (1 = black).
#include "iw_ctx.h"

#include <cstdlib>
#include <cstring>
#include <string>

namespace {

struct Reader {
	iw_context *ctx;
	iw_iodescr *io;
	bool byte(iw_byte *b)
	{
		size_t got = 0;
		if (!io->read_fn(ctx, io, b, 1, &got) || got != 1) {
			*b = 0;
			return false;
		}
		return true;
	}
	bool bytes(iw_byte *buf, size_t n)
	{
		size_t got = 0;
		return io->read_fn(ctx, io, buf, n, &got) && got == n;
	}
};

bool is_space(iw_byte b) { return b == 9 || b == 10 || b == 13 || b == 32; }

// Next whitespace-separated token of a PNM header; '#' comments run to EOL.
bool pnm_token(Reader &r, std::string &tok)
{
	tok.clear();
	bool comment = false;
	for (;;) {
		if (tok.size() >= 99) return false;
		iw_byte b;
		if (!r.byte(&b)) return false;
		if (comment) {
			if (b == 10) comment = false;
			continue;
		}
		if (b == '#') {
			comment = true;
			continue;
		}
		if (is_space(b)) {
			if (!tok.empty()) return true;
			continue;
		}
		tok.push_back((char)b);
	}
}

// One PAM header line (up to LF), truncated to 99 characters.
bool pam_line(Reader &r, std::string &line)
{
	line.clear();
	for (;;) {
		iw_byte b;
		if (!r.byte(&b)) return false;
		if (b == 0x0a) return true;
		if (line.size() < 99) line.push_back((char)b);
	}
}

}  // namespace

extern "C" int iw_read_pnm_file(struct iw_context *ctx, struct iw_iodescr *io)
{
	Reader r = {ctx, io};
	struct iw_image img;
	memset(&img, 0, sizeof(img));
	char sig[3];
	int code = 0, maxval = 0, nch = 0;
	std::string tok;
	bool header_ok = false;
	if (r.bytes((iw_byte *)sig, 3) && sig[0] == 'P') {
		if (sig[1] == '7' && sig[2] == 0x0a) code = 7;
		else if (sig[1] >= '1' && sig[1] <= '6' && is_space((iw_byte)sig[2])) code = sig[1] - '0';
	}
	if (!code) {
		iw_set_error(ctx, "Not a PNM/PAM file");
	} else if (code == 7) {
		std::string line;
		for (;;) {
			if (!pam_line(r, line)) break;
			if (!line.empty() && line[0] == '#') continue;
			size_t a = line.find_first_not_of(" \t\r\n");
			if (a == std::string::npos) continue;
			size_t b = line.find_first_of(" \t\r\n", a);
			std::string k = line.substr(a, b == std::string::npos ? std::string::npos : b - a);
			if (k == "ENDHDR") {
				header_ok = true;
				break;
			}
			std::string v;
			if (b != std::string::npos) {
				size_t c = line.find_first_not_of(" \t\r\n", b);
				if (c != std::string::npos) {
					size_t d = line.find_first_of(" \t\r\n", c);
					v = line.substr(c, d == std::string::npos ? std::string::npos : d - c);
				}
			}
			if (k == "WIDTH") img.width = atoi(v.c_str());
			else if (k == "HEIGHT") img.height = atoi(v.c_str());
			else if (k == "DEPTH") nch = atoi(v.c_str());
			else if (k == "MAXVAL") maxval = atoi(v.c_str());
		}
		if (header_ok && (maxval < 1 || maxval > 65535)) {
			iw_set_errorf(ctx, "Invalid max color value (%d)\n", maxval);
			header_ok = false;
		}
	} else if (code >= 4) {
		if (pnm_token(r, tok)) {
			img.width = atoi(tok.c_str());
			if (pnm_token(r, tok)) {
				img.height = atoi(tok.c_str());
				if (code == 4) {
					header_ok = true;
				} else if (pnm_token(r, tok)) {
					maxval = atoi(tok.c_str());
					if (maxval < 1 || maxval > 65535) iw_set_errorf(ctx, "Invalid max color value (%d)\n", maxval);
					else header_ok = true;
				}
			}
		}
	} else {
		iw_set_error(ctx, "Reading this PNM format is not supported");
	}
	if (!header_ok) {
		iw_set_error(ctx, "Error parsing header");
		return 0;
	}
	if (!iw_check_image_dimensions(ctx, img.width, img.height)) return 0;

	if (code == 5 || (code == 7 && nch == 1)) nch = 1;
	else if (code == 6 || (code == 7 && nch == 3)) nch = 3;
	else if (code == 7 && (nch == 2 || nch == 4)) {}
	else if (code != 4) {
		iw_set_error(ctx, "Unsupported PNM/PAM image type");
		return 0;
	}
	size_t bpr;
	if (code == 4) {
		img.imgtype = IW_IMGTYPE_GRAY;
		img.native_grayscale = 1;
		img.bit_depth = 1;
		bpr = (img.width + 7) / 8;
	} else {
		static const int types[5] = {0, IW_IMGTYPE_GRAY, IW_IMGTYPE_GRAYA, IW_IMGTYPE_RGB, IW_IMGTYPE_RGBA};
		img.imgtype = types[nch];
		if (nch == 1) img.native_grayscale = 1;
		img.bit_depth = (maxval >= 256) ? 16 : 8;
		bpr = (size_t)nch * (img.bit_depth / 8) * img.width;
		if (maxval != 255 && maxval != 65535)
			for (int c = 0; c < nch; c++) iw_set_input_max_color_code(ctx, c, maxval);
	}
	img.bpr = bpr;
	img.pixels = (iw_byte *)iw_malloc_large(ctx, bpr, img.height);
	if (!img.pixels) return 0;
	for (int j = 0; j < img.height; j++) {
		iw_byte *row = img.pixels + (size_t)j * bpr;
		if (!r.bytes(row, bpr)) {
			iw_free(ctx, img.pixels);
			return 0;
		}
		if (code == 4)
			for (size_t i = 0; i < bpr; i++) row[i] = (iw_byte)(255 - row[i]);
	}
	iw_set_input_image(ctx, &img);
	return 1;
}

extern "C" int iw_read_pam_file(struct iw_context *ctx, struct iw_iodescr *io)
{
	return iw_read_pnm_file(ctx, io);
}

namespace {

struct Writer {
	iw_context *ctx;
	iw_iodescr *io;
	void put(const void *buf, size_t n) { io->write_fn(ctx, io, buf, n); }
	void str(const std::string &s) { put(s.data(), s.size()); }
};

void pam_header(Writer &w, const struct iw_image &img, int nch, int maxval, const char *tupl)
{
	char buf[200];
	iw_snprintf(buf, sizeof(buf), "P7\nWIDTH %d\nHEIGHT %d\nDEPTH %d\nMAXVAL %d\nTUPLTYPE %s\nENDHDR\n",
		img.width, img.height, nch, maxval, tupl);
	w.str(buf);
}

bool bw_palette(iw_context *ctx, const struct iw_image &img)
{
	if (img.imgtype != IW_IMGTYPE_PALETTE) return false;
	if (!iw_get_value(ctx, IW_VAL_OUTPUT_PALETTE_GRAYSCALE)) return false;
	return iw_get_output_palette(ctx)->num_entries == 2;
}

// Max colour code for the written file: full range, or the reduced codes,
// which must agree across the channels the format stores.
bool file_maxval(iw_context *ctx, const struct iw_image &img, const int *cts, int n, int *maxval)
{
	*maxval = (img.bit_depth == 16) ? 65535 : 255;
	if (img.reduced_maxcolors) {
		*maxval = (int)img.maxcolorcode[cts[0]];
		for (int k = 1; k < n; k++) {
			if (img.maxcolorcode[cts[k]] != (unsigned int)*maxval) {
				iw_set_error(ctx, "PNM/PAM format requires equal bit depths");
				return false;
			}
		}
	}
	if (*maxval < 1 || *maxval > 65535) {
		iw_set_error(ctx, "Unsupported PNM/PAM bit depth");
		return false;
	}
	return true;
}

}  // namespace

extern "C" int iw_write_pnm_file(struct iw_context *ctx, struct iw_iodescr *io)
{
	struct iw_image img;
	iw_get_output_image(ctx, &img);
	Writer w = {ctx, io};
	const int req = iw_get_value(ctx, IW_VAL_OUTPUT_FORMAT);
	int fmt = req;
	if (req == IW_FORMAT_PNM) {
		if (img.imgtype == IW_IMGTYPE_PALETTE) fmt = IW_FORMAT_PBM;
		else if (img.imgtype == IW_IMGTYPE_GRAY) fmt = IW_FORMAT_PGM;
		else if (img.imgtype == IW_IMGTYPE_RGB) fmt = IW_FORMAT_PPM;
		else {
			iw_set_error(ctx, "Internal: Bad image type for PNM");
			return 0;
		}
	}
	if (img.bit_depth != 8 && img.bit_depth != 16 && img.imgtype != IW_IMGTYPE_PALETTE) return 0;
	const int bps = img.bit_depth / 8;
	char buf[200];
	if (fmt == IW_FORMAT_PBM) {
		if (!bw_palette(ctx, img)) {
			iw_set_error(ctx, "Cannot write this image type to a PBM file");
			return 0;
		}
		size_t rb = (img.width + 7) / 8;
		iw_snprintf(buf, sizeof(buf), "P4\n%d %d\n", img.width, img.height);
		w.str(buf);
		std::string row(rb, '\0');
		for (int j = 0; j < img.height; j++) {
			std::fill(row.begin(), row.end(), '\0');
			for (int i = 0; i < img.width; i++)
				if (img.pixels[(size_t)j * img.bpr + i] == 0) row[i / 8] |= (char)(1 << (7 - i % 8));
			w.put(row.data(), rb);
		}
		return 1;
	}
	const bool gray_like = img.imgtype == IW_IMGTYPE_GRAY || img.imgtype == IW_IMGTYPE_PALETTE;
	if (fmt == IW_FORMAT_PGM || (fmt == IW_FORMAT_PAM && gray_like)) {
		const bool bilevel = bw_palette(ctx, img);
		if (!bilevel && img.imgtype != IW_IMGTYPE_GRAY) {
			iw_set_error(ctx, "Cannot write non-grayscale image to PGM file");
			return 0;
		}
		int maxval = bilevel ? 1 : (img.bit_depth == 16 ? 65535 : 255);
		if (img.reduced_maxcolors) maxval = (int)img.maxcolorcode[IW_CHANNELTYPE_GRAY];
		if (maxval < 1 || maxval > 65535) {
			iw_set_error(ctx, "Unsupported PGM bit depth");
			return 0;
		}
		const size_t rb = (size_t)(bilevel ? 1 : bps) * img.width;
		if (fmt == IW_FORMAT_PAM) {
			pam_header(w, img, 1, maxval, maxval == 1 ? "BLACKANDWHITE" : "GRAYSCALE");
		} else {
			iw_snprintf(buf, sizeof(buf), "P5\n%d %d\n%d\n", img.width, img.height, maxval);
			w.str(buf);
		}
		for (int j = 0; j < img.height; j++) w.put(img.pixels + (size_t)j * img.bpr, rb);
		return 1;
	}
	if (fmt == IW_FORMAT_PAM && img.imgtype == IW_IMGTYPE_GRAYA) {
		static const int cts[2] = {IW_CHANNELTYPE_GRAY, IW_CHANNELTYPE_ALPHA};
		int maxval;
		if (!file_maxval(ctx, img, cts, 2, &maxval)) return 0;
		pam_header(w, img, 2, maxval, maxval == 1 ? "BLACKANDWHITE_ALPHA" : "GRAYSCALE_ALPHA");
		for (int j = 0; j < img.height; j++) w.put(img.pixels + (size_t)j * img.bpr, (size_t)2 * bps * img.width);
		return 1;
	}
	if (fmt == IW_FORMAT_PAM && img.imgtype == IW_IMGTYPE_RGBA) {
		static const int cts[4] = {IW_CHANNELTYPE_RED, IW_CHANNELTYPE_GREEN, IW_CHANNELTYPE_BLUE,
			IW_CHANNELTYPE_ALPHA};
		int maxval;
		if (!file_maxval(ctx, img, cts, 4, &maxval)) return 0;
		pam_header(w, img, 4, maxval, "RGB_ALPHA");
		for (int j = 0; j < img.height; j++) w.put(img.pixels + (size_t)j * img.bpr, (size_t)4 * bps * img.width);
		return 1;
	}
	if (fmt == IW_FORMAT_PPM || (fmt == IW_FORMAT_PAM && img.imgtype == IW_IMGTYPE_RGB)) {
		// PPM also takes grayscale images, expanded to RGB.
		static const int cts_rgb[3] = {IW_CHANNELTYPE_RED, IW_CHANNELTYPE_GREEN, IW_CHANNELTYPE_BLUE};
		static const int cts_g[1] = {IW_CHANNELTYPE_GRAY};
		const bool g = img.imgtype == IW_IMGTYPE_GRAY;
		int maxval;
		if (!file_maxval(ctx, img, g ? cts_g : cts_rgb, g ? 1 : 3, &maxval)) return 0;
		if (fmt == IW_FORMAT_PAM) pam_header(w, img, 3, maxval, "RGB");
		else {
			iw_snprintf(buf, sizeof(buf), "P6\n%d %d\n%d\n", img.width, img.height, maxval);
			w.str(buf);
		}
		std::string row((size_t)3 * bps * img.width, '\0');
		for (int j = 0; j < img.height; j++) {
			const iw_byte *src = img.pixels + (size_t)j * img.bpr;
			if (!g) {
				memcpy(&row[0], src, row.size());
			} else {
				for (int i = 0; i < img.width; i++)
					for (int k = 0; k < 3; k++)
						memcpy(&row[(size_t)(i * 3 + k) * bps], src + (size_t)i * bps, bps);
			}
			w.put(row.data(), row.size());
		}
		return 1;
	}
	iw_set_error(ctx, "Unsupported image type for PAM");
	return 0;
}

extern "C" int iw_write_pam_file(struct iw_context *ctx, struct iw_iodescr *io)
{
	return iw_write_pnm_file(ctx, io);
}

// Format dispatch (src/imagew-allfmts.c:18-182), for the formats this library
// carries.
extern "C" int iw_read_file_by_fmt(struct iw_context *ctx, struct iw_iodescr *io, int fmt)
{
	if (fmt == IW_FORMAT_PNM || fmt == IW_FORMAT_PAM) return iw_read_pnm_file(ctx, io);
	if (fmt == IW_FORMAT_MIFF) return iw_read_miff_file(ctx, io);
	const char *s = iw_get_fmt_name(fmt);
	if (!s) {
		iw_set_errorf(ctx, "Attempt to read unknown file format (%d)", fmt);
		return 0;
	}
	iw_set_errorf(ctx, "Reading %s files is not supported", s);
	return 0;
}

extern "C" int iw_write_file_by_fmt(struct iw_context *ctx, struct iw_iodescr *io, int fmt)
{
	iw_set_value(ctx, IW_VAL_OUTPUT_FORMAT, fmt);
	switch (fmt) {
	case IW_FORMAT_PNM:
	case IW_FORMAT_PPM:
	case IW_FORMAT_PGM:
	case IW_FORMAT_PBM:
	case IW_FORMAT_PAM:
		return iw_write_pnm_file(ctx, io);
	case IW_FORMAT_MIFF:
		return iw_write_miff_file(ctx, io);
	}
	const char *s = iw_get_fmt_name(fmt);
	if (!s) {
		iw_set_errorf(ctx, "Attempt to write unknown file format (%d)", fmt);
		return 0;
	}
	iw_set_errorf(ctx, "Writing %s files is not supported", s);
	return 0;
}
