// Internal state of an iw_context in the MI355X library.
//
// The fields mirror what the reference keeps in struct iw_context
// (src/imagew-internals.h:174-293) because the public setters and the
// decisions of iw_prepare_processing() read and write exactly that state;
// the pixel buffers of the reference (intermediate32, final_alpha32,
// dither_errors) have no host counterpart here -- they live in device memory
// for the duration of one iw_process_image() call.
#pragma once

#define IW_INCLUDE_UTIL_FUNCTIONS
#include "imagew.h"
#include "iwamd.h"

#include <cstddef>
#include <cstdint>
#include <vector>

#define IWX_MSG_MAX 200
#define IWX_DEFAULT_MAX_DIMENSION 40000      // src/imagew-internals.h:22
#define IWX_DEFAULT_MAX_MALLOC 2000000000u   // src/imagew-internals.h:23
#define IWX_NUM_CHANNELTYPES 5
#define IWX_CI_COUNT 4

#define IWX_BKGD_EARLY 1
#define IWX_BKGD_LATE 2

#define IWX_BKGD_SRC_NONE 0
#define IWX_BKGD_SRC_FILE 1
#define IWX_BKGD_SRC_REQ 2

struct iwx_resize_settings {
	int family = IW_RESIZETYPE_AUTO;
	int edge_policy = IW_EDGE_POLICY_STANDARD;
	int use_offset = 0;
	double param1 = 0.0, param2 = 0.0;
	double blur_factor = 1.0;
	double out_true_size = 0.0;
	double translate = 0.0;
	double channel_offset[3] = {0.0, 0.0, 0.0};
};

struct iwx_ci_in {
	int channeltype = 0;
	int disable_fast_get_sample = 0;
	double maxcolorcode_dbl = 0.0;
	int maxcolorcode_int = 0;
};

struct iwx_ci_intermed {
	int channeltype = 0;
	int cvt_to_grayscale = 0;
	int corresponding_input_channel = 0;
	int corresponding_output_channel = 0;
	double bkgd_color_lin = 0.0;
	int need_unassoc_alpha_processing = 0;
};

struct iwx_ci_out {
	int ditherfamily = 0, dithersubtype = 0, channeltype = 0;
	int color_count = 0;
	double maxcolorcode_dbl = 0.0;
	int maxcolorcode_int = 0;
	double bkgd1_color_lin = 0.0, bkgd2_color_lin = 0.0;
};

struct iwx_opt_ctx {
	int height = 0, width = 0, imgtype = 0, bit_depth = 0;
	size_t bpr = 0;
	const iw_byte *pixelsptr = nullptr;
	iw_byte *tmp_pixels = nullptr;
	int has_transparency = 0, has_partial_transparency = 0;
	int has_16bit_precision = 0, has_color = 0, palette_is_grayscale = 0;
	struct iw_palette *palette = nullptr;
	int has_colorkey_trns = 0;
	unsigned int colorkey[3] = {0, 0, 0};
	int has_bkgdlabel = 0;
	unsigned int bkgdlabel[4] = {0, 0, 0, 0};
};

struct iwx_option {
	char *name = nullptr;
	char *val = nullptr;
};

// Requests from the caller that may be adjusted before use
// (src/imagew-internals.h:129-172).
struct iwx_req {
	int output_depth = 0;
	int output_sample_type = 0;
	int output_maxcolorcode[IWX_NUM_CHANNELTYPES] = {0, 0, 0, 0, 0};
	int color_count[IWX_NUM_CHANNELTYPES] = {0, 0, 0, 0, 0};
	int out_true_valid = 0;
	double out_true_width = 0.0, out_true_height = 0.0;
	int output_rendering_intent = 0;
	int output_cs_valid = 0;
	struct iw_csdescr output_cs = {0, 0, 0.0};
	int suppress_output_cslabel = 0;
	int negate_target = 0;
	int bkgd_valid = 0;
	int bkgd_checkerboard = 0;
	struct iw_color bkgd = {{0, 0, 0, 0}};
	struct iw_color bkgd2 = {{0, 0, 0, 0}};
	int output_bkgd_label_valid = 0;
	struct iw_color output_bkgd_label = {{0, 0, 0, 0}};
	int use_bkgd_label_from_file = 0;
	int suppress_output_bkgd_label = 0;
	int output_format = 0, compression = 0, page_to_read = 0, include_screen = 1;
	int jpeg_samp_factor_h = 0, jpeg_samp_factor_v = 0, interlaced = 0;
	int bmp_no_fileheader = 0;
	std::vector<iwx_option> options;
};

struct iw_context {
	int caller_api_version = 0;
	int use_count = 0;
	unsigned int output_profile = 0;

	iw_mallocfn_type mallocfn = nullptr;
	iw_freefn_type freefn = nullptr;
	void *userdata = nullptr;
	iw_translatefn_type translate_fn = nullptr;
	iw_warningfn_type warning_fn = nullptr;

	iwx_ci_in img1_ci[IWX_CI_COUNT];
	struct iw_image img1 = {};
	struct iw_csdescr img1cs = {0, 0, 0.0};
	int img1_imgtype_logical = 0;
	int img1_numchannels_physical = 0, img1_numchannels_logical = 0;
	int img1_alpha_channel_index = 0;
	int img1_bkgd_label_set = 0;
	struct iw_color img1_bkgd_label_inputcs = {{0, 0, 0, 0}};
	struct iw_color img1_bkgd_label_lin = {{0, 0, 0, 0}};

	iwx_ci_intermed intermed_ci[IWX_CI_COUNT];
	int intermed_imgtype = 0, intermed_numchannels = 0, intermed_alpha_channel_index = 0;
	int intermed_canvas_width = 0, intermed_canvas_height = 0;

	struct iw_image img2 = {};
	struct iw_csdescr img2cs = {0, 0, 0.0};
	iwx_ci_out img2_ci[IWX_CI_COUNT];
	int img2_numchannels = 0;

	int ditherfamily_by_channeltype[IWX_NUM_CHANNELTYPES] = {0, 0, 0, 0, 0};
	int dithersubtype_by_channeltype[IWX_NUM_CHANNELTYPES] = {0, 0, 0, 0, 0};
	int uses_errdiffdither = 0;

	iwx_resize_settings resize_settings[2];
	int disable_rrctx_cache[2] = {0, 0};

	int to_grayscale = 0;
	int apply_bkgd = 0, apply_bkgd_strategy = 0, bkgd_checkerboard = 0;
	int bkgd_check_size = 0;
	int bkgd_check_origin[2] = {0, 0};
	int bkgd_color_source = 0;
	double bkgd1alpha = 0.0, bkgd2alpha = 0.0;

	int input_maxcolorcode_int = 0;
	double input_maxcolorcode = 0.0;
	int support_reduced_input_bitdepths = 0;
	int disable_output_lookup_tables = 0;
	int reduced_output_maxcolor_flag = 0;

	int randomize = 0, random_seed = 0;
	size_t max_malloc = IWX_DEFAULT_MAX_MALLOC;
	int max_width = IWX_DEFAULT_MAX_DIMENSION, max_height = IWX_DEFAULT_MAX_DIMENSION;

	int error_flag = 0;
	char *error_msg = nullptr;

	iwx_opt_ctx optctx;

	int no_gamma = 0, intclamp = 0, grayscale_formula = IW_GSF_STANDARD;
	double grayscale_weight[3] = {0, 0, 0};
	int pref_units = 0;
	iw_byte opt_grayscale = 1, opt_palette = 1, opt_16_to_8 = 1, opt_strip_alpha = 1,
		opt_binary_trns = 1;

	int canvas_width = 0, canvas_height = 0;
	int input_start_x = 0, input_start_y = 0, input_w = -1, input_h = -1;

	iwx_req req;
	struct iw_zlib_module *zlib_module = nullptr;

	// Library-specific state.
	int precision = IWAMD_PRECISION_PLAIN;
	int nearest_color_table_exists = 0;  // decided in prepare (src/imagew-main.c:1197-1236)
	int prepared = 0;                    // iwamd_prepare done, finish pending
	double last_timings_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	int num_last_timings = 0;
};

// iw_api.cpp
void iwx_warning(iw_context *ctx, const char *s);
void* iwx_malloc_large_check(iw_context *ctx, size_t n1, size_t n2);
int iwx_alloc_check(iw_context *ctx, size_t n1, size_t n2);

// iw_prepare.cpp
int iwx_prepare_processing(iw_context *ctx, int w, int h);
void iwx_fill_plan(iw_context *ctx, iwamd_plan *plan);
double iwx_x_to_linear(double v, const struct iw_csdescr *cs);
double iwx_linear_to_x(double v, const struct iw_csdescr *cs);
double iwx_color_to_grayscale(iw_context *ctx, double r, double g, double b);
struct iwx_opt_scan;
int iwx_finish_image(struct iw_context *ctx, const uint8_t *img2, const struct iwx_opt_scan *scan);

// What the output optimiser starts from, when the device has computed it over
// the finished image: the flags of src/imagew-opt.c:18-240, for each form a
// binary-transparency image can take the values of its varying colour-key
// component among opaque pixels (:418-683), and the set of distinct 8-bit
// colours while there are at most 256 (:719-797).  The device has also
// cleared the colour of fully transparent pixels (:1112-1163).
#define IWX_SCAN_COLOR   1u
#define IWX_SCAN_TRNS    2u
#define IWX_SCAN_PARTIAL 4u
#define IWX_SCAN_16BIT   8u
#define IWX_SCAN_HAVE_FLAGS 1u  // iwx_opt_scan::valid bits: what was computed
#define IWX_SCAN_HAVE_KEYS  2u
#define IWX_SCAN_HAVE_PAL   4u
#define IWX_SCAN_PAL_SLOTS 512
enum { IWX_KEY_RGB8 = 0, IWX_KEY_RGB16 = 1, IWX_KEY_GRAY8 = 2, IWX_KEY_GRAY16 = 3 };
struct iwx_opt_scan {
	unsigned int valid;       // IWX_SCAN_HAVE_*; 0: nothing (the host scans)
	unsigned int flags;       // IWX_SCAN_COLOR ...
	unsigned int used[4][8];  // [IWX_KEY_*]: 256-bit sets of the key component values in use
	unsigned int pal_count;   // colours in pal[]
	unsigned int pal_over;    // nonzero: more than 256 distinct colours
	// Open-addressed set of colours: 0 empty, else 2^32 | r | g<<8 | b<<16 | a<<24
	// (the colour of a fully transparent pixel taken as 0,0,0,0).
	unsigned long long pal[IWX_SCAN_PAL_SLOTS];
};

// iw_finish.cpp
int iwx_finish(iw_context *ctx, const iwx_opt_scan *scan = nullptr);
void iwx_optimize_image(iw_context *ctx, const iwx_opt_scan *scan = nullptr);

// iw_gpu.hip
int iwx_gpu_process(iw_context *ctx);
int iwx_gpu_device_count(void);
