// GPU executor of the ImageWorsener pixel pipeline for gfx950 (MI355X).
//
// One iw_process_image() resamples, then diffuses errors, on one HIP stream:
//   band_kernel  -- a band of output rows per workgroup: colour-in +
//                   premultiply/early background + vertical resample into LDS
//                   (FP32, as the reference stores it), then the horizontal
//                   resample and the per-pixel work of every channel (alpha
//                   first; unpremultiply, late background, quantise, ordered /
//                   random dither, store) (src/imagew-main.c:831-1155,
//                   669-765).  Error-diffusion channels stop after the clamp
//                   and hand their linear samples on.
//   vpass/hpass  -- the same work as two kernels with the intermediate in HBM,
//                   for rows too wide for a band in LDS.
//   fast_kernel  -- the chosen-precision form: both resamples as FP16 MFMA
//                   band products with FP32 sums, the per-pixel work in FP64.
//   errdiff      -- one workgroup per error-diffused plane walking the
//                   reference's serpentine order (src/imagew-main.c:497-567,
//                   1072-1136).
// Batches and the request loop run the same kernels over a descriptor table
// of many plans (persistent band kernels, hipGraph replay).
//
// Byte parity of the plain form with the reference: every product and sum is
// a separately rounded FP64 operation in the reference's order (contraction
// is off for this file and the compile line), no libm is called on the device
// (all transfer functions are host tables), and the intermediate and final
// alpha are rounded to FP32 where the reference stores them as float.
#include <hip/hip_runtime.h>

#include "iw_ctx.h"
#include "iw_tables.h"
#include "iw_server.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <tuple>
#include <type_traits>
#include <mutex>
#include <thread>
#include <vector>

#pragma clang fp contract(off)

namespace {

// Global-memory views of device pointers: accesses through them compile to
// global (not flat) loads and stores.  Flat accesses share the LDS counter,
// and with one outstanding every LDS wait in its reach becomes a full drain.
typedef __attribute__((address_space(1))) unsigned char g_u8;
typedef const __attribute__((address_space(1))) double g_cf64;
typedef const __attribute__((address_space(1))) unsigned short g_cu16;
typedef const __attribute__((address_space(1))) int g_cint;

// ---- device-side plan ----------------------------------------------------

struct DevWeights {
	int mode;            // WeightList::WEIGHTS / COPY
	const int *start;
	const int *src;
	const double *w;
	const int *copy_src;
	double edge;
	int edge_transparent;  // outside positions read edge (else the nearest sample)
	// The list by input position (iwx::DenseWeights).
	const int *dfirst, *dcount, *doff;
	const double *dw;
	int dmax;
};

struct DevIChan {
	int input_channel, cvt_gray, output_channel, need_unassoc, early_bkgd;
	double bkgd;
	const double *lut[3];
	double lut_lin_max;  // lut[0] is v / lut_lin_max (a linear channel), else 0
	DevWeights v, h;
};

enum { Q_THRESH = 0, Q_NC = 1 };

struct DevOChan {
	int family, subtype, qmode, P, is_alpha, ed_index;
	const double *T, *U, *lin, *code, *nc;
	const unsigned short *ix;          // QuantTable::ix (null: none)
	int ix_log2;
	const double *D;                   // QuantTable::D (null: not folded)
	int code_is_level;
	int ix_le2;                        // every bin of ix holds at most two entries
	const double *M;                   // error diffusion: round-up points (EdTables)
	int ed_warm;                       // error diffusion: warm-up pixels of a segment's walk
	double bk1, bk2;
	const uint64_t *prng_row, *prng_col;
};

struct DevPlan {
	const uint8_t *in;
	size_t in_bpr;
	int in_w, in_h, bit_depth, in_float, nch_phys;
	int in_f64;  // 32-bit float input held as linear doubles (converted on the host)
	unsigned int orient;
	int sx, sy, iw, ih;
	int alpha_in_index;
	const double *alpha_lut;  // v / alpha_max (linear)
	double alpha_max;
	int gsf;
	double gsw[3];
	int inter_nch, inter_alpha;
	int order[4];
	int W, Hp;
	DevIChan ic[4];
	float *inter;
	int intclamp;
	int out_w, out_h, out_depth, out_float, out_nch;
	size_t out_bpr;
	uint8_t *out;
	DevOChan oc[4];
	double *ed_samples;
	double *ed_err;
	int n_ed;
	int ed_seg;             // fewest pixels per error-diffusion segment
	// Per-request completion (request loop only; null otherwise): the last
	// error-diffusion plane of a request to finish copies the output image to
	// host memory and raises the flag there.
	int *ed_done;           // planes finished (device memory, zeroed before the run)
	uint8_t *host_out;      // device-visible view of pinned, coherent host memory
	int *host_flag;         // the same, set to 1 when host_out holds the output
	size_t out_bytes;
	int apply_bkgd, late, checker, check_size, ox, oy, bk_has_trans;
	int v_shared;  // every channel has the same vertical weights (dense form) in WEIGHTS mode
	int vband_r;   // two-pass form: output rows per thread of vband_kernel (0: vpass_kernel)
	int h_shared;  // the same for the horizontal weights
	double bk1a, bk2a;
	// Chosen-precision form: the shared horizontal list by 4-column blocks.
	// Output column i: (first block, block count, offset) in h4i[i]; block b
	// of it (columns 4b..4b+3) is h4w[4*(offset + b - first) + 0..3];
	// h4w[0..3] is a zero block.
	const int4 *h4i;
	const float *h4w;
};

// ---- input sampling (src/imagew-main.c:129-276) ---------------------------

__host__ __device__ inline void logical_to_physical(const DevPlan &P, int x, int y, int &rx, int &ry)
{
	const int X = P.sx + x, Y = P.sy + y;
	switch (P.orient) {
	case 0: rx = X; ry = Y; break;
	case 1: rx = P.in_w - 1 - X; ry = Y; break;
	case 2: rx = X; ry = P.in_h - 1 - Y; break;
	case 3: rx = P.in_w - 1 - X; ry = P.in_h - 1 - Y; break;
	case 4: rx = Y; ry = X; break;
	case 5: rx = Y; ry = P.in_w - 1 - X; break;
	case 6: rx = P.in_h - 1 - Y; ry = X; break;
	case 7: rx = P.in_h - 1 - Y; ry = P.in_w - 1 - X; break;
	default: rx = 0; ry = 0; break;
	}
}

__device__ __forceinline__ unsigned int raw_int(const DevPlan &P, int rx, int ry, int ch)
{
	const uint8_t *row = P.in + (size_t)ry * P.in_bpr;
	switch (P.bit_depth) {
	case 8: return row[P.nch_phys * rx + ch];
	case 16: {
		const uint8_t *q = row + (size_t)(P.nch_phys * rx + ch) * 2;
		return ((unsigned int)q[0] << 8) | q[1];
	}
	case 4: {
		unsigned int b = row[rx / 2];
		return (rx & 1) ? (b & 0x0f) : (b >> 4);
	}
	case 2: return (row[rx / 4] >> ((3 - rx % 4) * 2)) & 0x03;
	case 1: return (row[rx / 8] >> (7 - rx % 8)) & 0x01;
	}
	return 0;
}

__device__ __forceinline__ double raw_float(const DevPlan &P, int rx, int ry, int ch)
{
	if (P.bit_depth != 32) return 0.0;
	if (P.in_f64) {
		double d;
		memcpy(&d, P.in + (size_t)ry * P.in_bpr + (size_t)(P.nch_phys * rx + ch) * 8, 8);
		return d;
	}
	float f;
	memcpy(&f, P.in + (size_t)ry * P.in_bpr + (size_t)(P.nch_phys * rx + ch) * 4, 4);
	return (double)f;
}

// Sample k (0..2) of intermediate channel ic, in linear light.
__device__ __forceinline__ double sample_lin(const DevPlan &P, const DevIChan &ic, int k, int rx, int ry)
{
	const int ch = ic.input_channel + k;
	if (ch >= P.nch_phys) return 1.0;  // virtual alpha: opaque
	if (P.in_float) return raw_float(P, rx, ry, ch);
	return ic.lut[k][raw_int(P, rx, ry, ch)];
}

__device__ __forceinline__ double alpha_raw(const DevPlan &P, int rx, int ry)
{
	if (P.alpha_in_index >= P.nch_phys) return 1.0;
	if (P.in_float) return raw_float(P, rx, ry, P.alpha_in_index);
	return P.alpha_lut[raw_int(P, rx, ry, P.alpha_in_index)];
}

// ((w0*r + w1*g) + w2*b), optionally on sorted components
// (src/imagew-main.c:278-305).
__device__ __forceinline__ double to_gray(const DevPlan &P, double r, double g, double b)
{
	if (P.gsf == IW_GSF_WEIGHTED) return P.gsw[0] * r + P.gsw[1] * g + P.gsw[2] * b;
	double hi, mid, lo;
	if (g <= r) { hi = r; mid = g; } else { hi = g; mid = r; }
	if (b <= mid) {
		lo = b;
	} else {
		lo = mid;
		if (b <= hi) mid = b;
		else { mid = hi; hi = b; }
	}
	return P.gsw[0] * hi + P.gsw[1] * mid + P.gsw[2] * lo;
}

__device__ __forceinline__ double colour_in(const DevPlan &P, const DevIChan &ic, int x, int y)
{
	int rx, ry;
	logical_to_physical(P, x, y, rx, ry);
	double v;
	if (ic.cvt_gray)
		v = to_gray(P, sample_lin(P, ic, 0, rx, ry), sample_lin(P, ic, 1, rx, ry), sample_lin(P, ic, 2, rx, ry));
	else
		v = sample_lin(P, ic, 0, rx, ry);
	if (ic.need_unassoc) {
		v *= alpha_raw(P, rx, ry);
	} else if (ic.early_bkgd) {
		const double a = alpha_raw(P, rx, ry);
		v = a * v + (1.0 - a) * ic.bkgd;
	}
	return v;
}

__device__ __forceinline__ double clamp01(double v)
{
	if (v < 0.0) return 0.0;
	if (v > 1.0) return 1.0;
	return v;
}

// ---- vertical pass --------------------------------------------------------

__global__ void __launch_bounds__(256) vpass_kernel(DevPlan P)
{
	const int x = blockIdx.x * blockDim.x + threadIdx.x;
	const int j = blockIdx.y;
	if (x >= P.W) return;
	for (int c = 0; c < P.inter_nch; c++) {
		const DevIChan &ic = P.ic[c];
		double acc;
		if (ic.v.mode == iwx::WeightList::COPY) {
			const int s = ic.v.copy_src[j];
			acc = (s < 0) ? 0.0 : colour_in(P, ic, x, s);
		} else {
			acc = 0.0;
			const int t1 = ic.v.start[j + 1];
			for (int t = ic.v.start[j]; t < t1; t++) {
				const int s = ic.v.src[t];
				const double v = (s < 0) ? ic.v.edge : colour_in(P, ic, x, s);
				acc = acc + v * ic.v.w[t];
			}
		}
		if (P.intclamp) acc = clamp01(acc);
		P.inter[((size_t)c * P.Hp + j) * P.W + x] = (float)acc;
	}
}

// ---- quantisation and dithering (src/imagew-main.c:457-765) --------------

// Ordered-dither thresholds (k+0.5)/64 in the reference's two 8x8
// arrangements (src/imagew-main.c:460-480), stored as float like the
// reference's table.
__constant__ float c_dither_pattern[2][64] = {
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

// Number of entries of the ascending array a[0..n) that are <= s.  With the
// bin index ix (QuantTable::ix, 2^L bins of [0, 1]) a sample in [0, 1] needs
// only the entries of its own bin searched: the ix[b] before them lie below
// the bin's lower edge b/2^L <= s (s*2^L is exact), those from ix[b+1] on at
// or above its upper edge, which is > s.  Others (no index, s < 0, NaN) search
// the whole array.
__device__ __forceinline__ int qbin(int L, double s)
{
	const int NB = 1 << L;
	const double f = s * (double)NB;
	return f >= (double)(NB - 1) ? NB - 1 : (int)f;
}

__device__ __forceinline__ int count_le(const double *a_, int n, const unsigned short *ix_, int L, double s)
{
	const g_cf64 *a = (const g_cf64 *)a_;
	int lo = 0;
	if (ix_ && s >= 0.0) {
		const g_cu16 *ix = (const g_cu16 *)ix_;
		const int b = qbin(L, s);
		lo = ix[b];
		n = ix[b + 1] - lo;
	}
	while (n > 0) {
		const int half = n >> 1;
		if (a[lo + half] <= s) {
			lo += half + 1;
			n -= half + 1;
		} else {
			n = half;
		}
	}
	return lo;
}

// floor(e(s)): number of thresholds <= s.
__device__ __forceinline__ int level_floor(const DevOChan &q, double s)
{
	return count_le(q.T, q.P, (q.qmode == Q_THRESH && !q.D) ? q.ix : nullptr, q.ix_log2, s);
}

// Index of the first nearest-colour midpoint above s (0..255).
__device__ __forceinline__ int nc_search(const DevOChan &q, double s)
{
	return count_le(q.nc, 255, q.ix, q.ix_log2, s);
}

constexpr uint64_t kMontMinv = 0xfff0bf22ffffffffull;

// The reference's generator state for draw (i, j): Montgomery product of the
// row's Q and the column's b^-i, then x = (Q-1) / a.
__device__ __forceinline__ uint32_t prng_draw(const DevOChan &q, int i, int j)
{
	const uint64_t M = iwx::kPrngM;
	const uint64_t a = q.prng_row[j], b = q.prng_col[i];
	const uint64_t lo = a * b, hi = __umul64hi(a, b);
	const uint64_t minv = kMontMinv;  // -m^-1 mod 2^64, checked on the host
	const uint64_t u = lo * minv;
	const uint64_t ulo = u * M, uhi = __umul64hi(u, M);
	const uint64_t s_lo = lo + ulo;
	const uint64_t carry = (s_lo < lo) ? 1 : 0;
	uint64_t t = hi + uhi;
	const bool ov1 = t < hi;
	const uint64_t t2 = t + carry;
	const bool ov = ov1 || (t2 < t);
	t = t2;
	if (ov || t >= M) t -= M;
	return (uint32_t)((t - 1) / iwx::kPrngA);
}

// put_sample_convert_from_linear for the non-error-diffusion families.
__device__ __forceinline__ double quantize(const DevOChan &q, double s, int i, int j)
{
	if (q.qmode == Q_NC) return (double)nc_search(q, s);
	const g_cf64 *U = (const g_cf64 *)q.U, *lin = (const g_cf64 *)q.lin, *code = (const g_cf64 *)q.code;
	if (q.D) {
		// No dither, folded (QuantTable::D).
		const int k = count_le(q.D, q.P, q.ix, q.ix_log2, s);
		return q.code_is_level ? (double)k : code[k];
	}
	const int k = level_floor(q, s);
	if (k >= q.P || s < U[k]) return code[k];
	const double d_floor = s - lin[k];
	const double d_ceil = lin[k + 1] - s;
	switch (q.family) {
	case IW_DITHERFAMILY_NONE:
		return (d_ceil <= d_floor) ? code[k + 1] : code[k];
	case IW_DITHERFAMILY_ORDERED: {
		const double thr = (double)c_dither_pattern[q.subtype ? 1 : 0][(i % 8) + 8 * (j % 8)];
		return (d_floor / (d_floor + d_ceil) >= thr) ? code[k + 1] : code[k];
	}
	case IW_DITHERFAMILY_RANDOM: {
		const double thr = ((double)prng_draw(q, i, j)) / (double)0xffffffff;
		return (d_floor / (d_floor + d_ceil) >= thr) ? code[k + 1] : code[k];
	}
	}
	return 0.0;
}

// ---- horizontal pass ------------------------------------------------------

__device__ __forceinline__ double hgather(const DevPlan &P, int c, int j, int x)
{
	const DevIChan &ic = P.ic[c];
	const float *row = P.inter + ((size_t)c * P.Hp + j) * P.W;
	if (ic.h.mode == iwx::WeightList::COPY) {
		const int s = ic.h.copy_src[x];
		return (s < 0) ? 0.0 : (double)row[s];
	}
	double acc = 0.0;
	const int t1 = ic.h.start[x + 1];
	for (int t = ic.h.start[x]; t < t1; t++) {
		const int s = ic.h.src[t];
		const double v = (s < 0) ? ic.h.edge : (double)row[s];
		acc = acc + v * ic.h.w[t];
	}
	return acc;
}

// The plan fields the per-pixel work reads, copied out of the plan once per
// work item so that a loop over pixels keeps them in registers instead of
// loading them again for every pixel.  Slot idx is the idx-th channel in
// processing order (alpha first).
struct EmitCtx {
	int nch, intclamp, checker, check_size, ox, oy, bk_has_trans, late_bk, out_float, out_depth, out_nch;
	int out_w, out_h;
	double bk1a, bk2a;
	size_t out_bpr;
	uint8_t *out;
	double *ed_samples;
	int c[IWAMD_MAX_CH], o[IWAMD_MAX_CH], unassoc[IWAMD_MAX_CH], is_alpha[IWAMD_MAX_CH];
	DevOChan q[IWAMD_MAX_CH];
};

__device__ __forceinline__ void emit_ctx(const DevPlan &P, EmitCtx &E)
{
	E.nch = P.inter_nch;
	E.intclamp = P.intclamp;
	E.checker = P.checker;
	E.check_size = P.check_size;
	E.ox = P.ox;
	E.oy = P.oy;
	E.bk_has_trans = P.bk_has_trans;
	E.late_bk = P.apply_bkgd && P.late;
	E.out_float = P.out_float;
	E.out_depth = P.out_depth;
	E.out_nch = P.out_nch;
	E.out_w = P.out_w;
	E.out_h = P.out_h;
	E.bk1a = P.bk1a;
	E.bk2a = P.bk2a;
	E.out_bpr = P.out_bpr;
	E.out = P.out;
	E.ed_samples = P.ed_samples;
#pragma unroll
	for (int idx = 0; idx < IWAMD_MAX_CH; idx++) {
		const int c = idx < P.inter_nch ? P.order[idx] : 0;
		const DevIChan &ic = P.ic[c];
		E.c[idx] = c;
		E.o[idx] = idx < P.inter_nch ? ic.output_channel : -1;
		E.unassoc[idx] = ic.need_unassoc;
		E.is_alpha[idx] = c == P.inter_alpha;
		E.q[idx] = P.oc[E.o[idx] >= 0 ? E.o[idx] : 0];
	}
}

// The per-pixel work's plan fields for the common case -- integer output, no
// checkerboard or late background, every output channel either quantised
// without dither through one binned search (nearest-colour midpoints or
// QuantTable::D) or handed to error diffusion -- packed small so that a pixel
// loop keeps them in scalar registers.  Slot idx is the idx-th channel in
// processing order.  f: channel (bits 0-1), output channel + 1 (2-4; 0:
// none), unassociate (5), is alpha (6), error diffusion (7), code = count (8),
// bin-index log2 (9-13).  n: searched entries, or the plane index for error
// diffusion.
struct EmitSimple {
	int nch, intclamp, out_depth, out_nch, out_w, out_h;
	int pack4;  // 8-bit, four output channels, each written once, none diffused: one 32-bit store per pixel
	size_t out_bpr;
	uint8_t *out;
	double *ed_samples;
	int f[IWAMD_MAX_CH], n[IWAMD_MAX_CH];
	const double *sa[IWAMD_MAX_CH], *scode[IWAMD_MAX_CH];
	const unsigned short *six[IWAMD_MAX_CH];
};

__host__ __device__ inline bool emit_simple_ctx(const DevPlan &P, EmitSimple &E)
{
	if (P.out_float || P.checker || (P.apply_bkgd && P.late) || P.bk_has_trans) return false;
	E.nch = P.inter_nch;
	E.intclamp = P.intclamp;
	E.out_depth = P.out_depth;
	E.out_nch = P.out_nch;
	E.out_w = P.out_w;
	E.out_h = P.out_h;
	E.out_bpr = P.out_bpr;
	E.out = P.out;
	E.ed_samples = P.ed_samples;
	bool ok = true;
#pragma unroll
	for (int idx = 0; idx < IWAMD_MAX_CH; idx++) {
		const int c = idx < P.inter_nch ? P.order[idx] : 0;
		const DevIChan &ic = P.ic[c];
		const int o = idx < P.inter_nch ? ic.output_channel : -1;
		const DevOChan &q = P.oc[o >= 0 ? o : 0];
		const bool nc = q.qmode == Q_NC, ed = q.family == IW_DITHERFAMILY_ERRDIFF;
		E.sa[idx] = nc ? q.nc : q.D;
		E.scode[idx] = q.code;
		E.six[idx] = q.ix;
		E.n[idx] = ed ? q.ed_index : (nc ? 255 : q.P);
		E.f[idx] = c | ((o + 1) << 2) | (ic.need_unassoc ? 1 << 5 : 0) | (c == P.inter_alpha ? 1 << 6 : 0) |
			(ed ? 1 << 7 : 0) | ((nc || q.code_is_level) ? 1 << 8 : 0) | (q.ix_log2 << 9);
		if (o >= 0 && !ed && !(q.family == IW_DITHERFAMILY_NONE && q.ix && E.sa[idx] && E.n[idx] > 0)) ok = false;
	}
	// Rows of four-byte pixels start four-byte aligned (device blocks and
	// output arenas are 256-byte aligned).
	int seen = 0;
	bool pack = P.out_depth == 8 && P.out_nch == 4 && P.inter_nch == 4 && P.out_bpr % 4 == 0 &&
		((uintptr_t)P.out & 3) == 0;
#pragma unroll
	for (int idx = 0; idx < IWAMD_MAX_CH; idx++) {
		const int o = (E.f[idx] >> 2 & 7) - 1;
		if (o < 0 || (E.f[idx] & (1 << 7)) || (seen & (1 << o))) pack = false;
		else seen |= 1 << o;
	}
	E.pack4 = pack && seen == 15;
	return ok;
}

__device__ __forceinline__ void emit_store(const EmitCtx &E, int i, int j, int o, double s_full)
{
	g_u8 *row = (g_u8 *)E.out + (size_t)j * E.out_bpr;
	if (E.out_depth == 8) {
		row[E.out_nch * i + o] = (uint8_t)(0.5 + s_full);
	} else {
		const unsigned int v = (unsigned short)(0.5 + s_full);
		g_u8 *q = row + (size_t)(E.out_nch * i + o) * 2;
		q[0] = (uint8_t)(v >> 8);
		q[1] = (uint8_t)(v & 0xff);
	}
}

// One output pixel: every channel's horizontally resampled value (alpha
// first, rounded to FP32 like the reference's final_alpha32), unpremultiply,
// late background, then store, quantise, or hand the sample to error
// diffusion (src/imagew-main.c:1036-1123).  gather(c) returns the resampled
// value of intermediate channel c at this pixel.
template <class Gather>
__device__ __forceinline__ void emit_px(const EmitCtx &E, int i, int j, Gather gather)
{
	float alpha_f = 0.0f;
	int alt = 0;
	if (E.checker)
		alt = (((E.ox + i) / E.check_size) % 2) != (((E.oy + j) / E.check_size) % 2);
	const double bkalpha = E.bk_has_trans ? (alt ? E.bk2a : E.bk1a) : 0.0;
#pragma unroll
	for (int idx = 0; idx < IWAMD_MAX_CH; idx++) {
		if (idx >= E.nch) break;
		double v = gather(E.c[idx]);
		if (E.intclamp) v = clamp01(v);
		const bool is_alpha = E.is_alpha[idx];
		if (is_alpha) alpha_f = (float)v;
		const int o = E.o[idx];
		if (o < 0) continue;
		const DevOChan &q = E.q[idx];
		if (E.unassoc[idx]) {
			const double a = (double)alpha_f;
			if (a != 0.0) v /= a;
			if (E.late_bk) {
				const double bk = alt ? q.bk2 : q.bk1;
				if (E.bk_has_trans) v = v * a + bk * bkalpha * (1.0 - a);
				else v = v * a + bk * (1.0 - a);
			}
		} else if (is_alpha && E.bk_has_trans) {
			v = v + bkalpha * (1.0 - v);
		}
		if (E.out_float) {
			const float f = (float)v;
			memcpy(E.out + (size_t)j * E.out_bpr + (size_t)(E.out_nch * i + o) * 4, &f, 4);
			continue;
		}
		v = clamp01(v);
		if (q.family == IW_DITHERFAMILY_ERRDIFF) {
			((__attribute__((address_space(1))) double *)E.ed_samples)[((size_t)q.ed_index * E.out_h + j) * E.out_w + i] = v;
			continue;
		}
		emit_store(E, i, j, o, quantize(q, v, i, j));
	}
}

// emit_px for N pixels at once under EmitSimple.  The table searches of
// every channel and pixel go in stages -- all bin-index loads, then all
// entry loads, then the codes -- so their latencies overlap instead of
// adding up.  A bin holds at most two entries for 8-bit tables; a larger one
// (or a sample outside [0, 1]) is searched by count_le.  ALL: gather(n, g)
// fills g[c] with intermediate channel c of pixel n (c < nch) at once; else
// gather(n, c) returns one channel.  Pixels with live[n] false are computed,
// not stored.
template <int N, bool ALL, class Gather>
__device__ __forceinline__ void emit_px_simple(const EmitSimple &E, const int (&pi)[N], const int (&pj)[N],
	const bool (&live)[N], Gather gather)
{
	double v[IWAMD_MAX_CH][N];
	double g[N][IWAMD_MAX_CH];
	float alpha_f[N];
#pragma unroll
	for (int n = 0; n < N; n++) {
		alpha_f[n] = 0.0f;
		if constexpr (ALL) gather(n, g[n]);
	}
#pragma unroll
	for (int idx = 0; idx < IWAMD_MAX_CH; idx++) {
		if (idx >= E.nch) break;
		const int f = E.f[idx];
#pragma unroll
		for (int n = 0; n < N; n++) {
			double x;
			if constexpr (ALL) {
				x = g[n][0];
#pragma unroll
				for (int c = 1; c < IWAMD_MAX_CH; c++)
					if ((f & 3) == c) x = g[n][c];
			} else {
				x = gather(n, f & 3);
			}
			if (E.intclamp) x = clamp01(x);
			if (f & (1 << 6)) alpha_f[n] = (float)x;
			if (f & (1 << 5)) {
				const double a = (double)alpha_f[n];
				if (a != 0.0) x /= a;
			}
			v[idx][n] = clamp01(x);
		}
	}
	int lo[IWAMD_MAX_CH][N], hi[IWAMD_MAX_CH][N];
#pragma unroll
	for (int idx = 0; idx < IWAMD_MAX_CH; idx++) {
		if (idx >= E.nch) break;
		const int f = E.f[idx];
		if (!(f >> 2 & 7) || (f & (1 << 7))) continue;
		const g_cu16 *ix = (const g_cu16 *)E.six[idx];
#pragma unroll
		for (int n = 0; n < N; n++) {
			const int b = qbin(f >> 9 & 31, v[idx][n]);
			lo[idx][n] = ix[b];
			hi[idx][n] = ix[b + 1];
		}
	}
	double a0[IWAMD_MAX_CH][N], a1[IWAMD_MAX_CH][N];
#pragma unroll
	for (int idx = 0; idx < IWAMD_MAX_CH; idx++) {
		if (idx >= E.nch) break;
		const int f = E.f[idx];
		if (!(f >> 2 & 7) || (f & (1 << 7))) continue;
		const g_cf64 *a = (const g_cf64 *)E.sa[idx];
		const int last = E.n[idx] - 1;
#pragma unroll
		for (int n = 0; n < N; n++) {
			a0[idx][n] = a[min(lo[idx][n], last)];
			a1[idx][n] = a[min(lo[idx][n] + 1, last)];
		}
	}
	g_u8 *px[N];
	unsigned int word[N];
#pragma unroll
	for (int n = 0; n < N; n++) {
		px[n] = (g_u8 *)E.out + (size_t)pj[n] * E.out_bpr + (size_t)E.out_nch * pi[n] * (E.out_depth == 8 ? 1 : 2);
		word[n] = 0;
	}
#pragma unroll
	for (int idx = 0; idx < IWAMD_MAX_CH; idx++) {
		if (idx >= E.nch) break;
		const int f = E.f[idx], o = (f >> 2 & 7) - 1;
		if (o < 0) continue;
#pragma unroll
		for (int n = 0; n < N; n++) {
			if (!live[n]) continue;
			const double x = v[idx][n];
			if (f & (1 << 7)) {
				((__attribute__((address_space(1))) double *)E.ed_samples)[((size_t)E.n[idx] * E.out_h + pj[n]) *
					E.out_w + pi[n]] = x;
				continue;
			}
			const int l = lo[idx][n], m = hi[idx][n] - l;
			int k = l + (m > 0 && a0[idx][n] <= x) + (m > 1 && a1[idx][n] <= x);
			if (m > 2 || !(x >= 0.0)) k = count_le(E.sa[idx], E.n[idx], E.six[idx], f >> 9 & 31, x);
			const double code = (f & (1 << 8)) ? (double)k : ((const g_cf64 *)E.scode[idx])[k];
			if (E.pack4) {
				word[n] |= (unsigned int)(uint8_t)(0.5 + code) << (8 * o);
			} else if (E.out_depth == 8) {
				px[n][o] = (uint8_t)(0.5 + code);
			} else {
				const unsigned int u = (unsigned short)(0.5 + code);
				px[n][2 * o] = (uint8_t)(u >> 8);
				px[n][2 * o + 1] = (uint8_t)(u & 0xff);
			}
		}
	}
	if (E.pack4) {
#pragma unroll
		for (int n = 0; n < N; n++)
			if (live[n]) *(__attribute__((address_space(1))) unsigned int *)px[n] = word[n];
	}
}

template <class Gather>
__device__ __forceinline__ void emit_pixel(const DevPlan &P, int i, int j, Gather gather)
{
	EmitCtx E;
	emit_ctx(P, E);
	emit_px(E, i, j, gather);
}

__global__ void __launch_bounds__(256) hpass_kernel(DevPlan P)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const int j = blockIdx.y;
	if (i >= P.out_w) return;
	emit_pixel(P, i, j, [&](int c) { return hgather(P, c, j, i); });
}

// ---- fused band kernel ------------------------------------------------------
// One workgroup computes R consecutive output rows [j0, j0+R) completely: the
// vertical pass for those rows over every intermediate column into LDS (FP32,
// as the reference stores it), then the horizontal pass and the rest of the
// per-pixel work from LDS.  The intermediate never reaches HBM.
//
// Vertical pass by input position: each thread owns a column and keeps R
// accumulators per channel; it visits the band's input positions in
// ascending order, reads each input sample once, and adds v*w into every
// row's accumulator.  A row's weight list is in ascending position order
// (src/imagew-resize.c:282-327), so each accumulator receives its terms in the
// reference's order; positions a row does not list carry weight 0, which
// leaves its bits unchanged (an accumulator is never -0.0 and samples are
// finite).  In COPY mode (null / nearest) the row takes the sample itself.

#define BAND_MAX_K 256  // band input positions held in LDS

__device__ __forceinline__ double band_sample(const DevPlan &P, const DevIChan &ic, int x, int p, int num_in)
{
	if (p < 0 || p >= num_in) {
		// Only TRANSPARENT and REPLICATE list outside positions.
		if (ic.v.edge_transparent) return ic.v.edge;
		p = p < 0 ? 0 : num_in - 1;
	}
	return colour_in(P, ic, x, p);
}

// Direct colour-in for integer inputs of 8 or 16 bits whose intermediate
// channels each read one stored sample (no grey conversion, no virtual
// channel): the logical-to-physical map is affine, byte offset
// base + x*dx + y*dy for logical (x, y), and a pixel's alpha is looked up once
// for all its channels.  The arithmetic is colour_in's, in the same order.
// Every load is unconditional -- slots past the plan's channels repeat
// channel 0 and an unneeded alpha reads channel 0 too -- so a lane issues all
// of a row's loads before it waits for any.
struct FastCin {
	long long base, dx, dy;
	const double *lut[IWAMD_MAX_CH];
	double bkgd[IWAMD_MAX_CH], edge[IWAMD_MAX_CH];
	int chb[IWAMD_MAX_CH];   // byte of each channel's sample within a pixel
	int mode[IWAMD_MAX_CH];  // 0: as read, 1: times alpha, 2: early background
	int et[IWAMD_MAX_CH];    // rows outside the image read edge[c]
	int ab;                  // byte of the alpha sample
	const double *alut;      // its table (channel 0's when alpha is not needed)
	// 16-bit direct colour-in: a linear table's divisor (channel c, alpha),
	// 0 for the others.  A linear channel's value is (double)v / max, which
	// the device divides exactly (the table holds the same quotients).
	double lmax[IWAMD_MAX_CH], amax;
	int lean;                // no channel in mode 2 and none with a transparent edge
};

__host__ __device__ inline int fast_cin_depth(const DevPlan &P, FastCin &F)
{
	if (P.in_float || (P.bit_depth != 8 && P.bit_depth != 16)) return 0;
	const int bps = P.bit_depth / 8, pb = P.nch_phys * bps;
	bool need_a = false;
	for (int c = 0; c < IWAMD_MAX_CH; c++) {
		const DevIChan &ic = P.ic[c < P.inter_nch ? c : 0];
		if (ic.cvt_gray || ic.input_channel >= P.nch_phys || !ic.lut[0]) return 0;
		F.chb[c] = ic.input_channel * bps;
		F.lut[c] = ic.lut[0];
		F.mode[c] = ic.need_unassoc ? 1 : (ic.early_bkgd ? 2 : 0);
		F.bkgd[c] = ic.bkgd;
		F.edge[c] = ic.v.edge;
		F.et[c] = ic.v.edge_transparent;
		F.lmax[c] = ic.lut_lin_max;
		need_a |= F.mode[c] != 0;
	}
	F.lean = 1;
	for (int c = 0; c < IWAMD_MAX_CH; c++)
		if (F.mode[c] == 2 || F.et[c]) F.lean = 0;
	F.ab = F.chb[0];
	F.alut = F.lut[0];
	F.amax = F.lmax[0];
	if (need_a) {
		if (P.alpha_in_index >= P.nch_phys || !P.alpha_lut) return 0;
		F.ab = P.alpha_in_index * bps;
		F.alut = P.alpha_lut;
		F.amax = P.alpha_max;
	}
	int rx, ry;
	long long o[3];
	const int px[3] = {0, 1, 0}, py[3] = {0, 0, 1};
	for (int k = 0; k < 3; k++) {
		logical_to_physical(P, px[k], py[k], rx, ry);
		o[k] = (long long)ry * (long long)P.in_bpr + (long long)rx * pb;
	}
	F.base = o[0];
	F.dx = o[1] - o[0];
	F.dy = o[2] - o[0];
	// Four 8-bit samples per pixel, every pixel 4-byte aligned: one 32-bit
	// load per pixel ("depth" 32).  Four 16-bit samples, 8-byte aligned, lean:
	// one 64-bit load ("depth" 64).
	if (P.bit_depth == 8 && pb == 4 && ((F.base | F.dx | F.dy) & 3) == 0 && ((uintptr_t)P.in & 3) == 0) return 32;
	if (P.bit_depth == 16 && pb == 8 && F.lean && ((F.base | F.dx | F.dy) & 7) == 0 && ((uintptr_t)P.in & 7) == 0)
		return 64;
	return P.bit_depth;
}

// The plain form's vertical passes keep the tables for linear channels too:
// with its FP64 sums per band row, a division costs more there than a gather.
__host__ __device__ inline void fast_cin_tables_only(FastCin &F)
{
	for (int c = 0; c < IWAMD_MAX_CH; c++) F.lmax[c] = 0.0;
	F.amax = 0.0;
}

template <int DEPTH>
__device__ __forceinline__ unsigned int fast_raw(const g_u8 *in, long long off)
{
	if (DEPTH != 16) return in[off];
	return ((unsigned int)in[off] << 8) | in[off + 1];
}

// Intermediate channels 0..C-1 of logical input (xc, p), xc in [0, W): the
// band's sample, as band_sample(P, P.ic[c], xc, p, P.ih) would give it
// (channels from C on are left unspecified).  LEAN: no early background and
// no transparent edge (FastCin::lean), so a channel is its table value times
// alpha or times 1 (which leaves it unchanged).
template <int DEPTH, bool LEAN, class T = float>
__device__ __forceinline__ void fast_fetch(const DevPlan &P, const FastCin &F, long long cb, int p,
	T v[IWAMD_MAX_CH])
{
	const int C = P.inter_nch;
	if (DEPTH == 0) {
		// cb holds the column here.
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++)
			v[c] = c < C ? (T)band_sample(P, P.ic[c], (int)cb, p, P.ih) : (T)0;
		return;
	}
	const int pin = p < 0 ? 0 : (p >= P.ih ? P.ih - 1 : p);
	const long long off = cb + (long long)pin * F.dy;
	const g_u8 *in = (const g_u8 *)P.in;
	unsigned int raw[IWAMD_MAX_CH], araw;
	if (DEPTH == 32) {
		const unsigned int word = *(const __attribute__((address_space(1))) unsigned int *)(in + off);
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++) raw[c] = (word >> (8 * F.chb[c])) & 0xffu;
		araw = (word >> (8 * F.ab)) & 0xffu;
	} else if (DEPTH == 64) {
		// Big-endian 16-bit samples at even byte offsets of the pixel's word.
		const unsigned long long word = *(const __attribute__((address_space(1))) unsigned long long *)(in + off);
		auto be16 = [&](int b) {
			const unsigned int h = (unsigned int)(word >> (8 * b)) & 0xffffu;
			return ((h & 0xffu) << 8) | (h >> 8);
		};
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++) raw[c] = be16(F.chb[c]);
		araw = be16(F.ab);
	} else {
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++) raw[c] = fast_raw<DEPTH>(in, off + F.chb[c]);
		araw = fast_raw<DEPTH>(in, off + F.ab);
	}
	// 16-bit tables are 512 KiB, so a gather costs a cache line per lane; a
	// linear channel divides instead.
	auto value = [&](unsigned int v, const double *lut, double lmax) {
		if ((DEPTH == 16 || DEPTH == 64) && lmax > 0.0) return (double)v / lmax;
		return ((const g_cf64 *)lut)[v];
	};
	const double a = value(araw, F.alut, F.amax);
#pragma unroll
	for (int c = 0; c < IWAMD_MAX_CH; c++) {
		const double x = value(raw[c], F.lut[c], F.lmax[c]);
		if (LEAN) {
			v[c] = (T)(x * (F.mode[c] ? a : 1.0));
			continue;
		}
		const double xm = x * a;
		const double xb = a * x + (1.0 - a) * F.bkgd[c];
		double y = F.mode[c] == 1 ? xm : (F.mode[c] == 2 ? xb : x);
		if (p != pin && F.et[c]) y = F.edge[c];
		v[c] = (T)y;
	}
}

// The band's vertical pass when every channel shares one weight list
// (DevPlan::v_shared) and the input takes the direct colour-in: each column's
// pixels are read once for all channels, four input rows at a time with every
// load issued before the first is used.  Each accumulator still gets its terms
// in list order, with the same FP64 operations as colour_in and the generic
// pass.
// Columns x0, x0 + xstep, ... below W; row r of channel c of column x goes to
// out[c * cstride + r * W + x].
template <int R, int DEPTH, bool LEAN>
__device__ __forceinline__ void band_vertical_direct(const DevPlan &P, const FastCin &F, int j0, int rows,
	const double *wd, float *out, size_t cstride, int x0, int xstep)
{
	const int W = P.W, C = P.inter_nch;
	const DevWeights &dv = P.ic[0].v;
	int p_lo = 0x7fffffff, p_hi = -0x7fffffff;
	for (int r = 0; r < rows; r++) {
		const int n = dv.dcount[j0 + r];
		if (n == 0) continue;
		p_lo = min(p_lo, dv.dfirst[j0 + r]);
		p_hi = max(p_hi, dv.dfirst[j0 + r] + n - 1);
	}
	const int K = (p_hi >= p_lo) ? p_hi - p_lo + 1 : 0;
	for (int x = x0; x < W; x += xstep) {
		const long long cb = F.base + (long long)x * F.dx;
		double acc[IWAMD_MAX_CH][R];
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++)
#pragma unroll
			for (int r = 0; r < R; r++) acc[c][r] = 0.0;
		for (int k = 0; k < K; k += 4) {
			double v[4][IWAMD_MAX_CH];
#pragma unroll
			for (int t = 0; t < 4; t++) fast_fetch<DEPTH, LEAN, double>(P, F, cb, p_lo + min(k + t, K - 1), v[t]);
#pragma unroll
			for (int t = 0; t < 4; t++) {
				if (k + t >= K) break;
#pragma unroll
				for (int c = 0; c < IWAMD_MAX_CH; c++) {
					if (c >= C) break;
#pragma unroll
					for (int r = 0; r < R; r++) acc[c][r] = acc[c][r] + v[t][c] * wd[r * BAND_MAX_K + k + t];
				}
			}
		}
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++) {
			if (c >= C) break;
#pragma unroll
			for (int r = 0; r < R; r++) {
				if (r >= rows) break;
				double a = acc[c][r];
				if (P.intclamp) a = clamp01(a);
				out[c * cstride + (size_t)r * W + x] = (float)a;
			}
		}
	}
}

template <int R>
__device__ __forceinline__ void band_body(const DevPlan &P, int j0, unsigned char *band_lds)
{
	double *wd = (double *)band_lds;                       // [R][BAND_MAX_K]
	float *inter = (float *)(band_lds + sizeof(double) * R * BAND_MAX_K);  // [C][R][W]
	const int tid = threadIdx.x;
	const int rows = min(R, P.Hp - j0);
	const int W = P.W, C = P.inter_nch;
	const int num_in = P.ih;

	FastCin F;
	const int depth = P.v_shared ? fast_cin_depth(P, F) : 0;
	fast_cin_tables_only(F);
	for (int c = 0; c < (depth ? 1 : C); c++) {
		const DevIChan &ic = P.ic[c];
		const DevWeights &dv = ic.v;
		// The band's input positions.
		int p_lo = 0x7fffffff, p_hi = -0x7fffffff;
		for (int r = 0; r < rows; r++) {
			const int n = dv.dcount[j0 + r];
			if (n == 0) continue;
			p_lo = min(p_lo, dv.dfirst[j0 + r]);
			p_hi = max(p_hi, dv.dfirst[j0 + r] + n - 1);
		}
		const int K = (p_hi >= p_lo) ? p_hi - p_lo + 1 : 0;
		__syncthreads();  // previous channel's weights no longer read
		for (int t = tid; t < R * BAND_MAX_K; t += blockDim.x) {
			const int r = t / BAND_MAX_K, k = t % BAND_MAX_K;
			double w = 0.0;
			if (r < rows && k < K) {
				const int kk = p_lo + k - dv.dfirst[j0 + r];
				if (kk >= 0 && kk < dv.dcount[j0 + r]) w = dv.dw[dv.doff[j0 + r] + kk];
			}
			wd[t] = w;
		}
		__syncthreads();
		if (depth) {
			// Channel 0's weights are every channel's.
			const size_t cs = (size_t)R * W;
			const int x0 = threadIdx.x, xs = blockDim.x;
			if (depth == 32) {
				if (F.lean) band_vertical_direct<R, 32, true>(P, F, j0, rows, wd, inter, cs, x0, xs);
				else band_vertical_direct<R, 32, false>(P, F, j0, rows, wd, inter, cs, x0, xs);
			} else if (depth == 8) {
				if (F.lean) band_vertical_direct<R, 8, true>(P, F, j0, rows, wd, inter, cs, x0, xs);
				else band_vertical_direct<R, 8, false>(P, F, j0, rows, wd, inter, cs, x0, xs);
			} else if (depth == 64) {
				band_vertical_direct<R, 64, true>(P, F, j0, rows, wd, inter, cs, x0, xs);
			} else {
				if (F.lean) band_vertical_direct<R, 16, true>(P, F, j0, rows, wd, inter, cs, x0, xs);
				else band_vertical_direct<R, 16, false>(P, F, j0, rows, wd, inter, cs, x0, xs);
			}
			continue;
		}
		for (int x = tid; x < W; x += blockDim.x) {
			double acc[R];
#pragma unroll
			for (int r = 0; r < R; r++) acc[r] = 0.0;
			for (int k = 0; k < K; k++) {
				const double v = band_sample(P, ic, x, p_lo + k, num_in);
				if (dv.mode == iwx::WeightList::COPY) {
#pragma unroll
					for (int r = 0; r < R; r++)
						if (wd[r * BAND_MAX_K + k] != 0.0) acc[r] = v;
				} else {
#pragma unroll
					for (int r = 0; r < R; r++) acc[r] = acc[r] + v * wd[r * BAND_MAX_K + k];
				}
			}
#pragma unroll
			for (int r = 0; r < R; r++) {
				double a = acc[r];
				if (P.intclamp) a = clamp01(a);
				inter[((size_t)c * R + r) * W + x] = (float)a;
			}
		}
	}
	__syncthreads();
	// Horizontal pass and per-pixel work for the band's rows.
	const int n_px = P.out_w * rows;
	auto hsum = [&](int c, int r, int i) {
		const DevWeights &h = P.ic[c].h;
		const float *row = inter + ((size_t)c * R + r) * W;
		if (h.mode == iwx::WeightList::COPY) {
			const int s = h.copy_src[i];
			return (s < 0) ? 0.0 : (double)row[s];
		}
		const g_cint *start = (const g_cint *)h.start, *src = (const g_cint *)h.src;
		const g_cf64 *hw = (const g_cf64 *)h.w;
		double acc = 0.0;
		const int t1 = start[i + 1];
		for (int u = start[i]; u < t1; u++) {
			const int s = src[u];
			const double v = (s < 0) ? h.edge : (double)row[s];
			acc = acc + v * hw[u];
		}
		return acc;
	};
	EmitSimple ES;
	if (emit_simple_ctx(P, ES)) {
		// Every channel's horizontal sum of one pixel: through the shared dense
		// list (one weight load per tap for all channels) when the channels
		// share it, else channel by channel.  A dense position the list does
		// not hold weighs 0, which leaves a sum's bits unchanged.
		const DevWeights &h0 = P.ic[0].h;
		const g_cint *hfirst = (const g_cint *)h0.dfirst, *hcount = (const g_cint *)h0.dcount;
		const g_cint *hoff = (const g_cint *)h0.doff;
		const g_cf64 *hdw = (const g_cf64 *)h0.dw;
		auto hsum_all = [&](int r, int i, double (&g)[IWAMD_MAX_CH]) {
			if (!P.h_shared) {
#pragma unroll
				for (int c = 0; c < IWAMD_MAX_CH; c++) g[c] = c < C ? hsum(c, r, i) : 0.0;
				return;
			}
			double acc[IWAMD_MAX_CH] = {0.0, 0.0, 0.0, 0.0};
			const int f0 = hfirst[i], n = hcount[i], o0 = hoff[i];
			for (int u = 0; u < n; u++) {
				const double w = hdw[o0 + u];
				const int pos = f0 + u;
				const bool outside = pos < 0 || pos >= W;
				const int pc = pos < 0 ? 0 : (pos >= W ? W - 1 : pos);
#pragma unroll
				for (int c = 0; c < IWAMD_MAX_CH; c++) {
					if (c >= C) break;
					const DevWeights &h = P.ic[c].h;
					const double v = (outside && h.edge_transparent) ? h.edge :
						(double)inter[((size_t)c * R + r) * W + pc];
					acc[c] = acc[c] + v * w;
				}
			}
#pragma unroll
			for (int c = 0; c < IWAMD_MAX_CH; c++) g[c] = acc[c];
		};
		for (int t = tid; t < n_px; t += blockDim.x) {
			const int r[1] = {t / P.out_w}, pi[1] = {t % P.out_w};
			const int pj[1] = {j0 + r[0]};
			const bool live[1] = {true};
			emit_px_simple<1, true>(ES, pi, pj, live, [&](int n, double (&g)[IWAMD_MAX_CH]) { hsum_all(r[n], pi[n], g); });
		}
	} else {
		EmitCtx E;
		emit_ctx(P, E);
		for (int t = tid; t < n_px; t += blockDim.x) {
			const int r = t / P.out_w, i = t % P.out_w;
			emit_px(E, i, j0 + r, [&](int c) { return hsum(c, r, i); });
		}
	}
}

// Work item (plan, first row of the band).  Workgroup g first takes item g;
// with a counter (reset to 0 before the launch) it then keeps taking items
// gridDim.x + atomicAdd(counter, 1) until none is left -- a persistent
// kernel over a whole batch.  (The band lists keep this one queue: with the
// chosen form's per-XCD chunks the C3 batch's plain form ran 41.1 -> 45.3 ms.)
template <int R>
__global__ void __launch_bounds__(256) band_kernel(const DevPlan *__restrict__ plans,
	const int2 *__restrict__ items, int n_items, int *counter)
{
	extern __shared__ __align__(16) unsigned char band_lds[];
	__shared__ int next_item;
	int it = blockIdx.x;
	while (it < n_items) {
		const int2 w = items[it];
		band_body<R>(plans[w.x], w.y, band_lds);
		if (!counter) break;
		__syncthreads();  // LDS of this item is no longer read
		if (threadIdx.x == 0) next_item = (int)gridDim.x + atomicAdd(counter, 1);
		__syncthreads();
		it = next_item;
	}
}

// The two-pass form's vertical pass by bands of R intermediate rows, for
// plans whose channels share one vertical list (DevPlan::v_shared) and take
// the direct colour-in: each thread keeps one column's R rows per channel and
// visits the union of their input positions once, so every input sample is
// read and converted once per band instead of once per tap.  Each row's
// accumulator gets its list's terms in list order plus +0.0 for the band's
// other positions, which leaves it unchanged -- the same sums as vpass_kernel
// (src/imagew-main.c:831-927).  Grid: (ceil(W / 256), ceil(Hp / R)).
template <int R>
__global__ void __launch_bounds__(256) vband_kernel(DevPlan P)
{
	__shared__ double wd[R * BAND_MAX_K];
	const int j0 = blockIdx.y * R;
	const int rows = min(R, P.Hp - j0);
	const DevWeights &dv = P.ic[0].v;
	int p_lo = 0x7fffffff, p_hi = -0x7fffffff;
	for (int r = 0; r < rows; r++) {
		const int n = dv.dcount[j0 + r];
		if (n == 0) continue;
		p_lo = min(p_lo, dv.dfirst[j0 + r]);
		p_hi = max(p_hi, dv.dfirst[j0 + r] + n - 1);
	}
	const int K = (p_hi >= p_lo) ? p_hi - p_lo + 1 : 0;
	for (int t = threadIdx.x; t < R * BAND_MAX_K; t += blockDim.x) {
		const int r = t / BAND_MAX_K, k = t % BAND_MAX_K;
		double w = 0.0;
		if (r < rows && k < K) {
			const int kk = p_lo + k - dv.dfirst[j0 + r];
			if (kk >= 0 && kk < dv.dcount[j0 + r]) w = dv.dw[dv.doff[j0 + r] + kk];
		}
		wd[t] = w;
	}
	__syncthreads();
	FastCin F;
	const int depth = fast_cin_depth(P, F);
	fast_cin_tables_only(F);
	const int x0 = blockIdx.x * blockDim.x + threadIdx.x, xs = gridDim.x * blockDim.x;
	float *out = P.inter + (size_t)j0 * P.W;
	const size_t cs = (size_t)P.Hp * P.W;
	if (depth == 32) {
		if (F.lean) band_vertical_direct<R, 32, true>(P, F, j0, rows, wd, out, cs, x0, xs);
		else band_vertical_direct<R, 32, false>(P, F, j0, rows, wd, out, cs, x0, xs);
	} else if (depth == 8) {
		if (F.lean) band_vertical_direct<R, 8, true>(P, F, j0, rows, wd, out, cs, x0, xs);
		else band_vertical_direct<R, 8, false>(P, F, j0, rows, wd, out, cs, x0, xs);
	} else if (depth == 64) {
		band_vertical_direct<R, 64, true>(P, F, j0, rows, wd, out, cs, x0, xs);
	} else {
		if (F.lean) band_vertical_direct<R, 16, true>(P, F, j0, rows, wd, out, cs, x0, xs);
		else band_vertical_direct<R, 16, false>(P, F, j0, rows, wd, out, cs, x0, xs);
	}
}

// Queue the two-pass form's vertical pass of a plan: vband_kernel<vband_r>
// where the plan takes it, else vpass_kernel.
void enqueue_vpass(const DevPlan &D, hipStream_t st)
{
	const dim3 blk(256);
	switch (D.vband_r) {
	case 8: hipLaunchKernelGGL(vband_kernel<8>, dim3((D.W + 255) / 256, (D.Hp + 7) / 8), blk, 0, st, D); break;
	case 4: hipLaunchKernelGGL(vband_kernel<4>, dim3((D.W + 255) / 256, (D.Hp + 3) / 4), blk, 0, st, D); break;
	case 2: hipLaunchKernelGGL(vband_kernel<2>, dim3((D.W + 255) / 256, (D.Hp + 1) / 2), blk, 0, st, D); break;
	case 1: hipLaunchKernelGGL(vband_kernel<1>, dim3((D.W + 255) / 256, D.Hp), blk, 0, st, D); break;
	default: hipLaunchKernelGGL(vpass_kernel, dim3((D.W + 255) / 256, D.Hp), blk, 0, st, D); break;
	}
}

__global__ void reset_counters_kernel(int *c, int n)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) c[k] = 0;
}

// ---- chosen-precision form: FP32 MFMA, intermediate in registers ---------
//
// A band is FAST_ROWS (16) output rows; a tile is a band's rows x up to
// 16 * FAST_G output columns (one or two output groups of 16), and one wave
// takes a tile from input bytes to output bytes.  Both passes are dense
// products on v_mfma_f32_16x16x4f32 (FP32 operands and sums) with the band's
// weights zero outside each row's list:
//   vertical, transposed:  interT[x][r] = sum_k in[k][x] * Wv[r][k]
//     x: the 16 intermediate columns of a 16-aligned group; k: the band's
//     input rows, four per step;
//   horizontal:            out[r][i] = sum_x inter[r][x] * Wh[x][i]
//     i: the 16 output columns of an output group.
// Operand layout of v_mfma_f32_16x16x4f32 for lane l: A[l%16][l/16],
// B[l/16][l%16], D[4(l/16)+t][l%16] (t = 0..3).  The vertical product leaves
// lane l holding inter[r = l%16][x = xg + 4(l/16) + t] in D[t]: the A operand
// of the horizontal product's step t when that step sums the positions
// x = xg + 4k + t (k = 0..3 over the lanes' quarters).  So each intermediate
// group goes from the vertical accumulators into four horizontal steps
// without leaving the registers, and the horizontal accumulators leave lane l
// holding every channel of the pixels (row 4(l/16)+t, column i0 + l%16),
// which the per-pixel work takes from there.
//
// The grouping of the sums depends on the plan alone: intermediate groups
// are 16-aligned in absolute columns, an output group's horizontal steps run
// over the groups its own columns' lists reach (the same for both tile
// widths), and the vertical steps start at the band's first input row.  So a
// job's bytes depend neither on the tile width nor on the batch around it.
//
// A sample is its input table's value and premultiply (FP64, or in FP32
// through the band's FP32 copies of 8-bit tables) rounded to FP32; weights
// are the FP64 lists rounded to FP32; the vertical sum is FP32 where the
// reference rounds its intermediate to FP32.  Everything after the
// horizontal sum -- unpremultiply, background, the linear-to-output transfer,
// quantisation, dithering and error diffusion -- is the plain form's FP64
// code.

#define FAST_ROWS 16
#define FAST_MAX_KV 256   // input rows of a band (its weights live in LDS)
#define FAST_G 2          // output groups of 16 columns per tile, at most
#define FAST_MAX_TI (16 * FAST_G)
#define FAST_KC 5         // vertical steps whose input loads are issued together
#define FAST_RUN 64       // tiles per work item, at most
#define FAST_WG_PER_CU 4  // workgroups per CU of a persistent launch

// Kernel variants, by what a plan's tiles need (fast_variant): RGBA8 to RGBA8
// with the lean colour-in and straight-line per-pixel work (fast_rgba8_ok);
// 8-bit input through the LDS tables with the lean colour-in and the common
// per-pixel work; 16-bit input likewise; anything else.
enum { FAST_RGBA8 = 0, FAST_LEAN8 = 1, FAST_LEAN16 = 2, FAST_GENERIC = 3, FAST_NVAR = 4 };

typedef float float4_t __attribute__((ext_vector_type(4)));

struct FastItem {
	int job, j0, i0, ni;  // plan; first output row; first output column, count (<= FAST_MAX_TI)
	int x_lo, nx;         // intermediate columns the tile reads: 16-aligned start, count (multiple of 16)
	int p_lo, kv;         // the band's input rows
};

// Per workgroup: the band's vertical weights by input row (so a step's four
// rows x 16 output rows are 64 consecutive words) and FP32 copies of the
// 8-bit input tables.
#define FAST_QIX 2049  // bin index of an 8-bit quantiser table (2^11 bins)
struct __attribute__((aligned(16))) FastLds {
	float wv[FAST_MAX_KV][FAST_ROWS];
	float lut[IWAMD_MAX_CH + 1][256];  // channels, then alpha
	// FAST_RGBA8: the searched quantiser tables, colour then alpha.
	double qa[2][258];  // entries, then 2.0 (above any sample)
	unsigned short qix[2][FAST_QIX + 7];
};

// Whether a plan takes FAST_RGBA8's colour-in and per-pixel work: RGBA8
// input whose colour samples are premultiplied by alpha (byte 3, through its
// own table); the common per-pixel work (EmitSimple) with four channels
// processed alpha (channel 3) first, output channel c from intermediate
// channel c, 8-bit output, every colour channel unassociated, no clamp of
// the intermediate, no error diffusion; each channel's code the count of its
// table's entries <= s (levels as codes) over at most 256 entries with a
// 2^11-bin index whose bins hold at most two entries, one table for the three
// colour channels.
__host__ __device__ inline bool fast_rgba8_ok(const DevPlan &P, const EmitSimple &E, const FastCin &F)
{
	if (E.nch != 4 || P.inter_alpha != 3 || P.out_nch != 4 || P.out_depth != 8 || P.intclamp) return false;
	// Input: samples in bytes 0..3, alpha in byte 3 through channel 3's own
	// table; colour premultiplied, alpha not.
	for (int c = 0; c < 4; c++)
		if (F.chb[c] != c || F.mode[c] != (c < 3 ? 1 : 0)) return false;
	if (F.ab != 3 || F.alut != F.lut[3]) return false;
	// Byte offsets within the input fit 31 bits (fast_tile's row addressing).
	if ((double)P.in_bpr * (double)P.in_h >= 2147483648.0) return false;
	for (int o = 0; o < 4; o++)
		if (!P.oc[o].ix_le2) return false;
	for (int idx = 0; idx < 4; idx++) {
		const int f = E.f[idx], c = f & 3, o = (f >> 2 & 7) - 1;
		if (c != (idx + 3) % 4 || o != c || (f & (1 << 7)) || !(f & (1 << 8)) || (f >> 9 & 31) != 11) return false;
		if (((f >> 5) & 1) != (idx ? 1 : 0) || ((f >> 6) & 1) != (idx ? 0 : 1)) return false;
		if (E.n[idx] < 1 || E.n[idx] > 256 || !E.sa[idx] || !E.six[idx]) return false;
		if (idx > 1 && (E.sa[idx] != E.sa[1] || E.six[idx] != E.six[1] || E.n[idx] != E.n[1])) return false;
	}
	return true;
}

// FAST_RGBA8's code of a sample s in [0, 1] through LDS table q: the count of
// entries <= s.  The lo entries below s's bin are below s; the bin holds at
// most two more; every later entry (the table's own, or the 2.0 padding) is
// at or above the bin's upper edge, which is above s.
__device__ __forceinline__ int fast_q8(const FastLds &S, int q, double s)
{
	const int b = min((int)(s * 2048.0), 2047);
	const int lo = S.qix[q][b];
	return lo + (S.qa[q][lo] <= s) + (S.qa[q][lo + 1] <= s);
}

// 1/a for a finite nonzero double: the hardware reciprocal and one Newton
// step (within an ulp or so; the chosen form's sums carry FP32 rounding).
__device__ __forceinline__ double fast_rcp(double a)
{
	const double r = __builtin_amdgcn_rcp(a);
	return __builtin_fma(r, __builtin_fma(-a, r, 1.0), r);
}

// The 32-bit word of a 4-byte pixel: logical row p of the column at byte
// offset cb, clamped to the image.
__device__ __forceinline__ unsigned int fast_word(const DevPlan &P, const FastCin &F, long long cb, int p)
{
	const int pin = p < 0 ? 0 : (p >= P.ih ? P.ih - 1 : p);
	return *(const __attribute__((address_space(1))) unsigned int *)((const g_u8 *)P.in + cb + (long long)pin * F.dy);
}

// A loaded RGBA8-style word through the LDS tables (LEAN plans: a channel is
// its table value times alpha, or times 1).
__device__ __forceinline__ void fast_decode_word(const FastCin &F, const float (*lut)[256], unsigned int word,
	float v[IWAMD_MAX_CH])
{
	const float a = lut[IWAMD_MAX_CH][(word >> (8 * F.ab)) & 0xffu];
#pragma unroll
	for (int c = 0; c < IWAMD_MAX_CH; c++) v[c] = lut[c][(word >> (8 * F.chb[c])) & 0xffu] * (F.mode[c] ? a : 1.0f);
}

// fast_fetch for LEAN plans of 8-bit input through the LDS tables.
__device__ __forceinline__ void fast_fetch_lds8(const DevPlan &P, const FastCin &F, const float (*lut)[256],
	long long cb, int p, float v[IWAMD_MAX_CH])
{
	const int pin = p < 0 ? 0 : (p >= P.ih ? P.ih - 1 : p);
	const long long off = cb + (long long)pin * F.dy;
	const g_u8 *in = (const g_u8 *)P.in;
	unsigned int raw[IWAMD_MAX_CH];
#pragma unroll
	for (int c = 0; c < IWAMD_MAX_CH; c++) raw[c] = in[off + F.chb[c]];
	const float a = lut[IWAMD_MAX_CH][in[off + F.ab]];
#pragma unroll
	for (int c = 0; c < IWAMD_MAX_CH; c++) v[c] = lut[c][raw[c]] * (F.mode[c] ? a : 1.0f);
}

// fast_decode_word for FAST_RGBA8 plans: colour bytes 0..2 premultiplied by
// the alpha of byte 3, which is channel 3's value itself.
__device__ __forceinline__ void fast_decode_rgba8(const float (*lut)[256], unsigned int word, float v[IWAMD_MAX_CH])
{
	const float a = lut[3][word >> 24];
	v[0] = lut[0][word & 0xffu] * a;
	v[1] = lut[1][(word >> 8) & 0xffu] * a;
	v[2] = lut[2][(word >> 16) & 0xffu] * a;
	v[3] = a;
}

__device__ __forceinline__ float4_t mfma4(float a, float b, float4_t c)
{
	return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Floor division by 16 (columns can be negative: virtual pixels).
__device__ __forceinline__ int fast_grp(int x) { return x >> 4; }

// One tile on one wave.  rows: the band's output rows that exist.  SIMPLE:
// the plan takes the common per-pixel work (EmitSimple) -- known, so the
// general per-pixel code is left out; else simple says which applies.  RGBA8:
// the plan takes FAST_RGBA8's (fast_rgba8_ok; implies SIMPLE and LEAN).
template <int DEPTH, bool LEAN, bool SIMPLE, bool RGBA8 = false>
__device__ __forceinline__ void fast_tile(const DevPlan &P, const FastCin &F, const FastItem &w, const FastLds &S,
	const EmitSimple &ES, bool simple, int rows)
{
	const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
	const int C = P.inter_nch;
	const int kv4 = (w.kv + 3) & ~3;
	const int ng = (w.ni + 15) >> 4;
	const DevWeights &dh = P.ic[0].h;  // every channel shares the lists
	const g_cint *hfirst = (const g_cint *)dh.dfirst, *hcount = (const g_cint *)dh.dcount;
	const int4 *h4i = P.h4i;
	const __attribute__((address_space(1))) float4_t *h4w = (const __attribute__((address_space(1))) float4_t *)P.h4w;

	// Each output group: the lane's column's weights by 4-column blocks
	// (first block, block count, offset in h4w) and the intermediate groups
	// that the group's lists reach.
	int4 hb[FAST_G];
	int ga[FAST_G], gb[FAST_G];
	int g_first = 0x7fffffff, g_last = -0x7fffffff;
#pragma unroll
	for (int og = 0; og < FAST_G; og++) {
		const int ii = 16 * og + lr;
		int f = 0x7fffffff, e = -0x7fffffff;
		hb[og] = make_int4(0, 0, 0, 0);
		if (og < ng && ii < w.ni) {
			const int i = w.i0 + ii;
			hb[og] = h4i[i];
			const int n = hcount[i];
			if (n > 0) {
				f = hfirst[i];
				e = f + n - 1;
			}
		}
		for (int m = 1; m < 16; m <<= 1) {
			f = min(f, __shfl_xor(f, m));
			e = max(e, __shfl_xor(e, m));
		}
		// Every lane holds the same bounds now; say so to the compiler.
		ga[og] = __builtin_amdgcn_readfirstlane(e >= f ? fast_grp(f) : 0x7fffffff);
		gb[og] = __builtin_amdgcn_readfirstlane(e >= f ? fast_grp(e) : -0x7fffffff);
		if (og < ng && ga[og] <= gb[og]) {
			g_first = min(g_first, ga[og]);
			g_last = max(g_last, gb[og]);
		}
	}
	// Channels past the plan's are computed too in the lean forms (their
	// samples repeat channel 0, so they stay finite) and never read: the
	// products then run without a test between them.
	const int CM = SIMPLE && LEAN ? IWAMD_MAX_CH : C;
	bool fix = P.intclamp != 0;
#pragma unroll
	for (int c = 0; c < IWAMD_MAX_CH; c++) fix |= c < C && P.ic[c].h.edge_transparent;

	float4_t acc[FAST_G][IWAMD_MAX_CH];
#pragma unroll
	for (int og = 0; og < FAST_G; og++)
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++) acc[og][c] = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};

	// FAST_RGBA8, rows inside the image, ascending addresses: each load is a
	// uniform row address plus the lane's 32-bit byte offset (the image is
	// under 2^31 bytes, fast_rgba8_ok), the lane's part of it computed once
	// per group; only the steps the band has are loaded, and a group's first
	// FAST_KC steps are issued before the previous group's horizontal
	// products (in flight meanwhile).  (Bands that reach past an edge take
	// the clamped loads below: a clamped form of these loads spilled.)
	bool direct = false;
	const g_u8 *row0 = nullptr;
	if constexpr (RGBA8 && DEPTH == 32) {
		direct = w.p_lo >= 0 && w.p_lo + kv4 <= P.ih && F.dx > 0 && F.dy > 0;
		row0 = (const g_u8 *)P.in + F.base + (long long)w.p_lo * F.dy;
	}
	const unsigned int dy = (unsigned int)F.dy;
	auto group_off = [&](int g) {
		const int xq = 16 * g + lr;
		const int xcq = xq < 0 ? 0 : (xq >= P.W ? P.W - 1 : xq);
		return (unsigned int)xcq * (unsigned int)F.dx + (unsigned int)lq * dy;
	};
	unsigned int pre[FAST_KC];
	auto issue = [&](unsigned int vo, int k0) {
#pragma unroll
		for (int s = 0; s < FAST_KC; s++)
			if (k0 + 4 * s < kv4)
				pre[s] = *(const __attribute__((address_space(1))) unsigned int *)(row0 +
					(size_t)(vo + (unsigned int)(k0 + 4 * s) * dy));
		// Keeps the loads here, all in flight together.
		asm volatile("" ::: "memory");
	};
	unsigned int vo_next = 0;  // the lane's offset in the group whose loads are issued
	if (direct) {
		vo_next = group_off(g_first);
		issue(vo_next, 0);
	}

	for (int g = g_first; g <= g_last; g++) {
		const int xg = 16 * g;
		// This group's horizontal weights, loaded before the vertical steps.
		float4_t hw[FAST_G];
#pragma unroll
		for (int og = 0; og < FAST_G; og++) {
			const int rel = 4 * g + lq - hb[og].x;
			hw[og] = h4w[(rel >= 0 && rel < hb[og].y) ? hb[og].z + rel : 0];
		}
		const int x = xg + lr;
		const int xc = x < 0 ? 0 : (x >= P.W ? P.W - 1 : x);
		const long long cb = DEPTH ? F.base + (long long)xc * F.dx : (long long)xc;
		float4_t dv[IWAMD_MAX_CH];
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++) dv[c] = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
		if (direct) {
			const unsigned int vo = vo_next;
			for (int k0 = 0; k0 < kv4; k0 += 4 * FAST_KC) {
				if (k0 > 0) issue(vo, k0);
#pragma unroll
				for (int s = 0; s < FAST_KC; s++) {
					if (k0 + 4 * s >= kv4) break;
					float v[IWAMD_MAX_CH];
					fast_decode_rgba8(S.lut, pre[s], v);
					const float b = S.wv[k0 + 4 * s + lq][lr];
#pragma unroll
					for (int c = 0; c < IWAMD_MAX_CH; c++) dv[c] = mfma4(v[c], b, dv[c]);
				}
			}
			if (g < g_last) {
				vo_next = group_off(g + 1);
				issue(vo_next, 0);
			}
		} else if constexpr (DEPTH == 32 && LEAN) {
			// FAST_KC steps' words are loaded before the first is used.
			for (int k0 = 0; k0 < kv4; k0 += 4 * FAST_KC) {
				unsigned int wd[FAST_KC];
#pragma unroll
				for (int s = 0; s < FAST_KC; s++) wd[s] = fast_word(P, F, cb, w.p_lo + k0 + 4 * s + lq);
				// Keeps the loads here, all in flight together (otherwise each is
				// sunk to its use behind the step's test and waited for alone).
				asm volatile("" ::: "memory");
#pragma unroll
				for (int s = 0; s < FAST_KC; s++) {
					if (k0 + 4 * s >= kv4) break;
					float v[IWAMD_MAX_CH];
					fast_decode_word(F, S.lut, wd[s], v);
					const float b = S.wv[k0 + 4 * s + lq][lr];
#pragma unroll
					for (int c = 0; c < IWAMD_MAX_CH; c++)
						if (c < CM) dv[c] = mfma4(v[c], b, dv[c]);
				}
			}
		} else {
			for (int k = 0; k < kv4; k += 4) {
				float v[IWAMD_MAX_CH];
				if constexpr (LEAN && DEPTH == 8)
					fast_fetch_lds8(P, F, S.lut, cb, w.p_lo + k + lq, v);
				else
					fast_fetch<DEPTH, LEAN>(P, F, cb, w.p_lo + k + lq, v);
				const float b = S.wv[k + lq][lr];
#pragma unroll
				for (int c = 0; c < IWAMD_MAX_CH; c++)
					if (c < CM) dv[c] = mfma4(v[c], b, dv[c]);
			}
		}
		// The intermediate: a virtual column outside a TRANSPARENT edge is the
		// edge value; intclamp clamps.
#pragma unroll
		for (int c = 0; c < IWAMD_MAX_CH; c++) {
			if (!fix || c >= C) continue;
#pragma unroll
			for (int t = 0; t < 4; t++) {
				float v = dv[c][t];
				const int xt = xg + 4 * lq + t;
				if ((xt < 0 || xt >= P.W) && P.ic[c].h.edge_transparent) v = (float)P.ic[c].h.edge;
				if (P.intclamp) v = fminf(fmaxf(v, 0.0f), 1.0f);
				dv[c][t] = v;
			}
		}
#pragma unroll
		for (int og = 0; og < FAST_G; og++) {
			if (og >= ng || g < ga[og] || g > gb[og]) continue;
#pragma unroll
			for (int t = 0; t < 4; t++)
#pragma unroll
				for (int c = 0; c < IWAMD_MAX_CH; c++)
					if (c < CM) acc[og][c] = mfma4(dv[c][t], hw[og][t], acc[og][c]);
		}
	}

	if constexpr (RGBA8) {
		// Alpha first: its code, and alpha rounded to FP32 (as the reference's
		// final_alpha32) whose reciprocal unassociates the colour channels
		// (one FP64 division per pixel; the sums themselves carry FP32
		// rounding already).  Then one 32-bit store per pixel.
		const bool pack = ES.pack4;
		auto clamp = [](double v) { return fmin(fmax(v, 0.0), 1.0); };
		// Stores: the band's first output row (uniform) plus the lane's 32-bit
		// byte offset within the band's 16 rows.
		g_u8 *const orow = (g_u8 *)P.out + (size_t)w.j0 * P.out_bpr;
		const unsigned int obpr = (unsigned int)P.out_bpr, lrow = (unsigned int)(4 * lq) * obpr;
#pragma unroll
		for (int og = 0; og < FAST_G; og++) {
			if (og >= ng) break;
			const int ii = 16 * og + lr, i = w.i0 + ii;
#pragma unroll
			for (int t = 0; t < 4; t++) {
				const int r = 4 * lq + t;
				const float af = acc[og][3][t];
				const double a = (double)af;
				unsigned int word = (unsigned int)fast_q8(S, 1, clamp(a)) << 24;
				const double inv = af != 0.0f ? fast_rcp(a) : 1.0;
#pragma unroll
				for (int c = 0; c < 3; c++)
					word |= (unsigned int)fast_q8(S, 0, clamp((double)acc[og][c][t] * inv)) << (8 * c);
				if (ii < w.ni && r < rows) {
					g_u8 *px = orow + (size_t)(lrow + (unsigned int)t * obpr + 4u * (unsigned int)i);
					if (pack) {
						*(__attribute__((address_space(1))) unsigned int *)px = word;
					} else {
#pragma unroll
						for (int c = 0; c < 4; c++) px[c] = (uint8_t)(word >> (8 * c));
					}
				}
			}
		}
		return;
	}
	// Per-pixel work in FP64, as in the plain form: one pixel per step, the
	// step's sums picked out of the accumulators (og and t are uniform), so
	// that the per-pixel code appears once.
	auto pick = [&](int og, int t, int c) {
		const float4_t &a = og ? acc[FAST_G - 1][c] : acc[0][c];
		return t == 0 ? a[0] : (t == 1 ? a[1] : (t == 2 ? a[2] : a[3]));
	};
#pragma unroll 1
	for (int e = 0; e < 4 * ng; e++) {
		const int og = e >> 2, t = e & 3;
		const int ii = 16 * og + lr, i = w.i0 + ii, r = 4 * lq + t;
		const bool live = ii < w.ni && r < rows;
		if (SIMPLE || simple) {
			const int pi[1] = {i}, pj[1] = {w.j0 + r};
			const bool lv[1] = {live};
			// The flags are made opaque per step: otherwise every condition
			// derived from them is computed once and kept live across the
			// tile as a lane mask, which spills.
			EmitSimple E = ES;
#pragma unroll
			for (int c = 0; c < IWAMD_MAX_CH; c++) {
				asm volatile("" : "+s"(E.f[c]));
				asm volatile("" : "+s"(E.n[c]));
			}
			emit_px_simple<1, true>(E, pi, pj, lv, [&](int, double (&gv)[IWAMD_MAX_CH]) {
#pragma unroll
				for (int c = 0; c < IWAMD_MAX_CH; c++) gv[c] = (double)pick(og, t, c);
			});
		} else if constexpr (!SIMPLE) {
			if (live) {
				EmitCtx E;
				emit_ctx(P, E);
				emit_px(E, i, w.j0 + r, [&](int c) {
					return (double)(c == 0 ? pick(og, t, 0) : c == 1 ? pick(og, t, 1) : c == 2 ? pick(og, t, 2) : pick(og, t, 3));
				});
			}
		}
	}
}

// A work item: tiles tiles[first .. first+count) of one band.  The band's
// weights and the input tables go to LDS once; then the waves take the tiles
// in turn.
template <int VAR>
__device__ __forceinline__ void fast_band(const DevPlan &P, const FastItem *__restrict__ tiles, int first, int count,
	FastLds &S)
{
	// The wave index is uniform: so is every tile field read with it, which
	// keeps the tiles' loops free of lane masks.
	const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const FastItem w0 = tiles[first];
	const int kv4 = (w0.kv + 3) & ~3;
	const DevWeights &dv = P.ic[0].v;
	{
		const g_cint *vfirst = (const g_cint *)dv.dfirst, *vcount = (const g_cint *)dv.dcount;
		const g_cint *voff = (const g_cint *)dv.doff;
		const g_cf64 *vw = (const g_cf64 *)dv.dw;
		for (int t = tid; t < kv4 * FAST_ROWS; t += blockDim.x) {
			const int k = t >> 4, r = t & 15;
			float v = 0.0f;
			const int j = w0.j0 + r;
			if (j < P.Hp && k < w0.kv) {
				const int kk = w0.p_lo + k - vfirst[j];
				if (kk >= 0 && kk < vcount[j]) v = (float)vw[voff[j] + kk];
			}
			S.wv[k][r] = v;
		}
	}
	FastCin F;
	const int depth = fast_cin_depth(P, F);
	if ((depth == 8 || depth == 32) && F.lean)
		for (int t = tid; t < 256; t += blockDim.x) {
#pragma unroll
			for (int c = 0; c < IWAMD_MAX_CH; c++) S.lut[c][t] = (float)((const g_cf64 *)F.lut[c])[t];
			S.lut[IWAMD_MAX_CH][t] = (float)((const g_cf64 *)F.alut)[t];
		}
	EmitSimple ES;
	const bool simple = emit_simple_ctx(P, ES);
	if constexpr (VAR == FAST_RGBA8) {
		// The colour table (slot 0, from the first colour channel) and the
		// alpha table (slot 1).
		for (int t = tid; t < 2 * 258; t += blockDim.x) {
			const int q = t >= 258, e = t - q * 258, idx = q ? 0 : 1;
			S.qa[q][e] = e < ES.n[idx] ? ((const g_cf64 *)ES.sa[idx])[e] : 2.0;
		}
		for (int t = tid; t < 2 * FAST_QIX; t += blockDim.x) {
			const int q = t >= FAST_QIX, e = t - q * FAST_QIX, idx = q ? 0 : 1;
			S.qix[q][e] = ((const g_cu16 *)ES.six[idx])[e];
		}
	}
	__syncthreads();
	const int rows = min(FAST_ROWS, P.Hp - w0.j0);
	for (int tt = wave; tt < count; tt += (int)(blockDim.x >> 6)) {
		const FastItem w = tiles[first + tt];
		if constexpr (VAR == FAST_RGBA8) {
			if (depth == 32) fast_tile<32, true, true, true>(P, F, w, S, ES, true, rows);
			else fast_tile<8, true, true, true>(P, F, w, S, ES, true, rows);
		} else if constexpr (VAR == FAST_LEAN8) {
			if (depth == 32) fast_tile<32, true, true>(P, F, w, S, ES, true, rows);
			else fast_tile<8, true, true>(P, F, w, S, ES, true, rows);
		} else if constexpr (VAR == FAST_LEAN16) {
			if (depth == 64) fast_tile<64, true, true>(P, F, w, S, ES, true, rows);
			else fast_tile<16, true, true>(P, F, w, S, ES, true, rows);
		} else {
			switch (depth) {
			case 32:
				if (F.lean) fast_tile<32, true, false>(P, F, w, S, ES, simple, rows);
				else fast_tile<32, false, false>(P, F, w, S, ES, simple, rows);
				break;
			case 8:
				if (F.lean) fast_tile<8, true, false>(P, F, w, S, ES, simple, rows);
				else fast_tile<8, false, false>(P, F, w, S, ES, simple, rows);
				break;
			case 64: fast_tile<64, true, false>(P, F, w, S, ES, simple, rows); break;
			case 16:
				if (F.lean) fast_tile<16, true, false>(P, F, w, S, ES, simple, rows);
				else fast_tile<16, false, false>(P, F, w, S, ES, simple, rows);
				break;
			default: fast_tile<0, false, false>(P, F, w, S, ES, simple, rows); break;
			}
		}
	}
	__syncthreads();  // the next item's weights replace these
}

// The chosen form's work loop.  Without counters workgroup g runs item g.
// With counters (XCD_COUNT of them, reset to 0 before the launch) the items
// are XCD_COUNT contiguous chunks, chunk[x] .. chunk[x + 1] - 1, and
// workgroup g takes the items of chunk g % XCD_COUNT one at a time, then
// those left in the other chunks.  Workgroups are dispatched round-robin
// over the device's XCDs, so a chunk -- neighbouring bands of the same
// images -- runs on one XCD, and the input rows that neighbouring bands
// share are read through that XCD's L2 once (C3 batch: DRAM fetch 2.1x ->
// 1.2x the input).  Two barriers per take: every thread has finished the
// previous item (its LDS is free) and has read the new index before the
// next take.
#define XCD_COUNT 8

// The next item for a workgroup with counters: x is the chunk it takes
// from, d how many chunks it has found empty; -1 when all are.
__device__ __forceinline__ int xcd_take(int *counter, const int *__restrict__ chunk, int &x, int &d, int &next_item)
{
	while (d < XCD_COUNT) {
		if (threadIdx.x == 0) next_item = chunk[x] + atomicAdd(counter + x, 1);
		__syncthreads();
		const int it = next_item;
		__syncthreads();
		if (it < chunk[x + 1]) return it;
		d++;
		x = (x + 1) % XCD_COUNT;
	}
	return -1;
}

// Work item (first tile, tile count).  VAR: the variant every item's plan
// has (fast_variant).
template <int VAR>
__global__ void __launch_bounds__(256, VAR == FAST_GENERIC ? 1 : FAST_WG_PER_CU) fast_kernel(
	const DevPlan *__restrict__ plans, const FastItem *__restrict__ tiles, const int2 *__restrict__ items,
	int n_items, int *counter, const int *__restrict__ chunk)
{
	__shared__ FastLds S;
	__shared__ int next_item;
	int x = (int)(blockIdx.x % XCD_COUNT), d = 0;
	int it = counter ? xcd_take(counter, chunk, x, d, next_item) : ((int)blockIdx.x < n_items ? (int)blockIdx.x : -1);
	while (it >= 0) {
		const int2 band = items[it];  // first tile, tile count
		fast_band<VAR>(plans[tiles[band.x].job], tiles, band.x, band.y, S);
		it = counter ? xcd_take(counter, chunk, x, d, next_item) : -1;
	}
}

// ---- error diffusion ------------------------------------------------------
//
// A plane is one dependency chain: rows alternate direction (src/imagew-main.c:
// 1072-1077), the sample of each pixel takes the errors of the one or two
// pixels before it in its row (src/imagew-main.c:497-567), and a row needs
// the errors of the rows above it.  Within a row the chain is cut by
// speculation that is exact by construction:
//   - the row is split into segments, one per lane; each lane walks its
//     segment starting `warm` pixels early with no in-row carry;
//   - what the rest of a row depends on at a pixel is the error of the one or
//     two pixels before it (the walk's state).  Once a speculative walk's
//     state equals the true chain's at some pixel, bit for bit, every later
//     pixel of it is the true one.  Two walks that take the same levels see
//     their difference shrink by the in-row weights at every step (7/16 for
//     Floyd-Steinberg), so their sums round to the same doubles within a few
//     dozen pixels;
//   - segment k's state at its first pixel is compared with segment k-1's
//     after its last; where they differ, one lane walks the true chain on
//     from the last verified segment until its state equals the speculative
//     walk's, overwriting what it walked, and the checks resume from there.
// Each row then leaves its errors for the next two rows as an ordered
// gather (stage A): for each target cell, the row's sources in its scan
// order, so the same terms are added in the same order as the reference's
// per-pixel scatter.  An exact pixel (no error in the reference) carries
// +0.0: adding a signed zero to an accumulator that is never -0.0 leaves it
// unchanged, so the bits agree without flags.
//
// The decision at a pixel is the reference's floor/ceil choice
// (src/imagew-main.c:712-748) through tables built on the host (no libm on
// the device): level k = #{T <= s}, found through a 4096-bin index and, in
// the common case, one compare; exact below U[k] (error 0); else rounded up
// from M[k], the least s where (lin[k+1]-s) <= (s-lin[k]) holds in FP64;
// error s - lin[level].

// Error-diffusion weights; slots 0,1: x+1,x+2 of this row; 2..6: x-2..x+2 of
// the next row; 7..11: the row after (src/imagew-main.c:510-535).
__constant__ double c_ed_matrix[8][12] = {
	{7.0/16, 0.0, 0.0, 3.0/16, 5.0/16, 1.0/16, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},          // Floyd-Steinberg
	{7.0/48, 5.0/48, 3.0/48, 5.0/48, 7.0/48, 5.0/48, 3.0/48, 1.0/48, 3.0/48, 5.0/48, 3.0/48, 1.0/48},  // JJN
	{8.0/42, 4.0/42, 2.0/42, 4.0/42, 8.0/42, 4.0/42, 2.0/42, 1.0/42, 2.0/42, 4.0/42, 2.0/42, 1.0/42},  // Stucki
	{8.0/32, 4.0/32, 2.0/32, 4.0/32, 8.0/32, 4.0/32, 2.0/32, 0.0, 0.0, 0.0, 0.0, 0.0},  // Burkes
	{5.0/32, 3.0/32, 2.0/32, 4.0/32, 5.0/32, 4.0/32, 2.0/32, 0.0, 2.0/32, 3.0/32, 2.0/32, 0.0},  // Sierra3
	{4.0/16, 3.0/16, 1.0/16, 2.0/16, 3.0/16, 2.0/16, 1.0/16, 0.0, 0.0, 0.0, 0.0, 0.0},  // Sierra2
	{2.0/4, 0.0, 0.0, 1.0/4, 1.0/4, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},                 // Sierra 4-2-1
	{1.0/8, 1.0/8, 0.0, 1.0/8, 1.0/8, 1.0/8, 0.0, 0.0, 0.0, 1.0/8, 0.0, 0.0}};          // Atkinson

#define ED_WALKERS 256    // waves 0-3: up to one segment per thread
#define ED_THREADS 512    // waves 4-7 help: next row's samples, previous row's codes
#define ED_LDS_TABLE 256  // largest level count - 1 whose tables live in LDS
#define ED_ROWS_LDS 3072  // widest row whose working rows live in LDS
#define ED_NB 4096        // bins of the level index (s * 4096 is exact)
#define ED_SEG_MIN 8      // fewest pixels per segment (default)

// How a row of W pixels is cut: nseg segments of seg pixels (the last one
// shorter).  Rows are stored transposed -- pixel z (scan order) at
// (z % seg) * nsp + z / seg -- so that the segments' walks, which advance in
// step, read and write consecutive elements; nsp (odd) spreads the row
// stage's strided accesses over the LDS banks.
struct EdCut {
	int nseg, seg, nsp;
	float inv;  // 1 / seg: (z + 0.5) * inv truncates to z / seg exactly for the row widths here
};

__host__ __device__ inline EdCut ed_cut(int W, int seg_min)
{
	EdCut c;
	int n = W / seg_min;
	n = n < 1 ? 1 : (n > ED_WALKERS ? ED_WALKERS : n);
	c.seg = (W + n - 1) / n;
	c.nseg = (W + c.seg - 1) / c.seg;
	c.nsp = c.nseg | 1;
	c.inv = 1.0f / (float)c.seg;
	return c;
}

// Elements of a transposed row (> W): seg * nsp positions, then a guard
// element at index seg * nsp.  The guard's (sample, carried error) pair is
// (0, 0), which decides to level 0 with error 0 (level 0 is 0.0 in linear
// light): the warm-up walks read it for pixels before the row's start.
__host__ __device__ inline int ed_cap(int W, int seg_min)
{
	const EdCut c = ed_cut(W, seg_min);
	return c.seg * c.nsp + 1;
}

// Per-plane working rows in HBM, offsets in doubles, for rows wider than
// ED_ROWS_LDS: R1 and the next row's samples (W each, by x and by scan
// position), then transposed: the errors, (sample, carried error) pairs
// (16-byte aligned) and two level rows (u16).  total is even, so every
// plane's rows start 16-byte aligned.
struct EdRowsOff {
	size_t bnext, err, be, lvl0, lvl1, total;
};

__host__ __device__ inline EdRowsOff ed_rows_off(int W, int seg_min)
{
	const size_t cap = (size_t)ed_cap(W, seg_min), l4 = (cap + 3) / 4;
	EdRowsOff o;
	o.bnext = (size_t)W;
	o.err = 2 * (size_t)W;
	o.be = o.err + cap + (cap & 1);
	o.lvl0 = o.be + 2 * cap;
	o.lvl1 = o.lvl0 + l4;
	o.total = o.lvl1 + l4;
	o.total += o.total & 1;
	return o;
}

// Level k's decision data as one record: T = the threshold at which level
// k+1 starts (2.0 past the last level), U, M and lin as in the tables.
struct __align__(16) EdRec {
	double T, U, M, lin;
};

// Number of thresholds T[k], k in [a, b), that are <= s, plus a.
__host__ __device__ inline int level_in(const double *T, int a, int b, double s)
{
	int n = b - a;
	while (n > 0) {
		const int half = n >> 1;
		if (T[a + half] <= s) {
			a += half + 1;
			n -= half + 1;
		} else {
			n = half;
		}
	}
	return a;
}

// A channel with at most ED_SMALL_P + 1 levels keeps its level records in
// registers: T[k] where level k+1 starts (2.0 past the last level), U, M
// and lin as in the tables (lin[P+1] = lin[P]).
#define ED_SMALL_P 3
struct EdSmall {
	double T[ED_SMALL_P], U[ED_SMALL_P + 1], M[ED_SMALL_P + 1], lin[ED_SMALL_P + 2];
};

// The tables one plane's walk reads.  With the records in LDS (up to 257
// levels) the common decision reads two bin bounds, then three records;
// with at most ED_SMALL_P + 1 levels it compares in registers; else the
// level comes from a bisection over the global thresholds.
struct EdDecide {
	const EdRec *rec;             // levels 0..P and two guards (LDS)
	const unsigned short *Q;      // Q[q] = level of q / ED_NB, q = 0..ED_NB
	const double *T, *U, *M, *lin;  // global tables (no records)
	int P;
	EdSmall sm;                   // the records of a small channel
};

// The reference's decision for a clamped sample s through EdSmall: the level
// is the count of thresholds <= s (they ascend, so the last one passed names
// the level's record), then as ed_decide.
__host__ __device__ inline double ed_decide_small(const EdSmall &m, double s, int &level)
{
	int k = 0;
	double U = m.U[0], M = m.M[0], lk = m.lin[0], lk1 = m.lin[1];
#pragma unroll
	for (int i = 0; i < ED_SMALL_P; i++) {
		const bool up = s >= m.T[i];
		k += up ? 1 : 0;
		U = up ? m.U[i + 1] : U;
		M = up ? m.M[i + 1] : M;
		lk = up ? m.lin[i + 1] : lk;
		lk1 = up ? m.lin[i + 2] : lk1;
	}
	const bool exact = s < U, ceil_ = s >= M;
	level = (exact || !ceil_) ? k : k + 1;
	return exact ? 0.0 : s - (ceil_ ? lk1 : lk);
}

// A channel's records in EdSmall (P <= ED_SMALL_P) from its global tables.
__host__ __device__ inline void ed_small_fill(EdSmall &m, int P, const double *T, const double *U, const double *M,
	const double *lin)
{
	for (int i = 0; i < ED_SMALL_P; i++) m.T[i] = i < P ? T[i] : 2.0;
	for (int i = 0; i <= ED_SMALL_P; i++) {
		m.U[i] = i <= P ? U[i] : 2.0;
		m.M[i] = i <= P ? M[i] : 2.0;
	}
	for (int i = 0; i <= ED_SMALL_P + 1; i++) m.lin[i] = lin[i <= P ? i : P];
}

// The reference's decision for a clamped sample s: its error (0.0 when
// exact) and output level.
template <bool TAB>
__host__ __device__ inline double ed_decide(const EdDecide &d, double s, int &level)
{
	const int q = (int)(s * ED_NB) < ED_NB - 1 ? (int)(s * ED_NB) : ED_NB - 1;
	const int k0 = d.Q[q], k1 = d.Q[q + 1];
	int k;
	double U, M, lk, lk1;
	if (TAB) {
		// The common bin holds at most one threshold, T[k0]; a bin with more
		// (levels narrower than a bin) bisects its thresholds afterwards.
		const EdRec r0 = d.rec[k0], r1 = d.rec[k0 + 1];
		const double l2 = d.rec[k0 + 2].lin;
		const bool up = s >= r0.T;
		k = k0 + (up ? 1 : 0);
		U = up ? r1.U : r0.U;
		M = up ? r1.M : r0.M;
		lk = up ? r1.lin : r0.lin;
		lk1 = up ? l2 : r1.lin;
		if (__builtin_expect(k1 > k0 + 1, 0)) {
			int a = k0, n = k1 - k0;
			while (n > 0) {
				const int half = n >> 1;
				if (d.rec[a + half].T <= s) {
					a += half + 1;
					n -= half + 1;
				} else {
					n = half;
				}
			}
			k = a;
			U = d.rec[a].U;
			M = d.rec[a].M;
			lk = d.rec[a].lin;
			lk1 = d.rec[a + 1].lin;
		}
	} else {
		k = (k1 == k0) ? k0 : level_in(d.T, k0, k1, s);
		U = d.U[k];  // U[P] = +inf: the top level is exact
		M = d.M[k];  // M[P] = 2.0: never rounds up
		lk = d.lin[k];
		lk1 = d.lin[k < d.P ? k + 1 : d.P];
	}
	const bool exact = s < U, ceil_ = s >= M;
	level = (exact || !ceil_) ? k : k + 1;
	return exact ? 0.0 : s - (ceil_ ? lk1 : lk);
}

// One step of a walk: the pixel's (sample, error from earlier rows) and the
// errors of the one or two pixels before it in its row -> its error and
// level.  The carried terms are added in the reference's order, x-2 before
// x-1 (separately rounded FP64 operations); the clamp is exact (the sum is
// never NaN or -0.0).
// TABM: 0 global tables (bisection), 1 LDS records, 2 EdSmall in registers.
template <int TABM, bool HAS_M1>
__device__ __forceinline__ double ed_step(const EdDecide &d, double2 be, double p1, double p2, double m0, double m1,
	int &level)
{
	const double cell = HAS_M1 ? (be.y + p2 * m1) + p1 * m0 : be.y + p1 * m0;
	const double s = fmin(fmax(be.x + cell, 0.0), 1.0);
	if constexpr (TABM == 2) return ed_decide_small(d.sm, s, level);
	else return ed_decide<TABM == 1>(d, s, level);
}

#define ED_CAP_LDS (ED_ROWS_LDS + 32)  // >= ed_cap(W) for W <= ED_ROWS_LDS

template <int TABM, bool ROWS>
struct EdLds {
	EdRec rec[TABM == 1 ? ED_LDS_TABLE + 3 : 1];   // levels 0..P, two guards
	unsigned short Q[TABM == 2 ? 1 : ED_NB + 2];
	double2 be[ROWS ? ED_CAP_LDS : 1];             // (sample, carried error), transposed
	double err[ROWS ? ED_CAP_LDS : 1];             // the row's errors, transposed
	unsigned short lvl[2][ROWS ? ED_CAP_LDS : 1];  // levels of rows j (j % 2), transposed
	double bnext[ROWS ? ED_ROWS_LDS : 1];          // the next row's samples by scan position
	double R1[ROWS ? ED_ROWS_LDS : 1];             // row j-1's share for row j+1, by x
	double sw[ED_WALKERS][2], se[ED_WALKERS][2];   // segment states: at its first pixel, after its last
	unsigned long long fail[ED_WALKERS / 64];      // boundaries whose states differ
};

struct EdCtx {
	EdDecide d;
	g_cf64 *code;              // output code per level
	g_cf64 *samp;              // the plane's clamped samples [H][W]
	double *R1, *err, *bnext;
	double2 *be;
	unsigned short *lvl0, *lvl1;  // level rows (HBM working rows only)
	double m0, m1;
	int W, o;
	EdCut cut;
	g_u8 *out;
	size_t out_bpr;
	int out_nch, out_depth;
};

__device__ __forceinline__ int ed_idx(const EdCut &k, int z)
{
	const int q = (int)(((float)z + 0.5f) * k.inv);
	return (z - q * k.seg) * k.nsp + q;
}

// emit_store() through the context's global view of the output.
__device__ __forceinline__ void ed_store(const EdCtx &c, int i, int j, double s_full)
{
	g_u8 *row = c.out + (size_t)j * c.out_bpr;
	if (c.out_depth == 8) {
		row[c.out_nch * i + c.o] = (uint8_t)(0.5 + s_full);
	} else {
		const unsigned int v = (unsigned short)(0.5 + s_full);
		g_u8 *q = row + (size_t)(c.out_nch * i + c.o) * 2;
		q[0] = (uint8_t)(v >> 8);
		q[1] = (uint8_t)(v & 0xff);
	}
}

// Row j's codes from its levels lv, threads t, t + nt, ...
__device__ __forceinline__ void ed_store_row(const EdCtx &c, const unsigned short *lv, int j, int t, int nt)
{
	const int W = c.W;
	const bool rev = (j & 1) != 0;
	for (int z = t; z < W; z += nt) ed_store(c, rev ? W - 1 - z : z, j, c.code[lv[ed_idx(c.cut, z)]]);
}

// Row j's samples by scan position, threads [t0, t0 + nt).
__device__ __forceinline__ void ed_load_row(const EdCtx &c, int j, int t, int nt)
{
	const int W = c.W;
	const bool rev = (j & 1) != 0;
	const g_cf64 *row = c.samp + (size_t)j * W;
	for (int z = t; z < W; z += nt) c.bnext[z] = row[rev ? W - 1 - z : z];
}

// Stage A of row j, all threads: each pixel's sample (prefetched) and the
// error carried to it from the rows above -- what row j-2 left in R1, then
// row j-1's sources x-2f, x-f, x, x+f, x+2f in that row's scan order with
// weights m[6..2] (the reference's scatter, src/imagew-main.c:497-567, as an
// ordered gather); R1[x] then takes row j-1's share for row j+1 (weights
// m[11..7]).  Row j-1 runs the other way, so pixel z of row j is its scan
// position W-1-z, and its sources are that position's neighbours.
__device__ __forceinline__ void ed_stage_a(const EdCtx &c, const double *m, int j)
{
	const int W = c.W;
	const bool rev = (j & 1) != 0;
	for (int z = threadIdx.x; z < W; z += ED_THREADS) {
		const int x = rev ? W - 1 - z : z;
		double e = 0.0;
		if (j > 0) {
			const int zp = W - 1 - z;
			double a2 = 0.0;
			e = c.R1[x];
			for (int dd = -2; dd <= 2; dd++) {
				const int p = zp + dd;
				if (p < 0 || p >= W) continue;
				const double er = c.err[ed_idx(c.cut, p)];
				e = e + er * m[4 - dd];
				a2 = a2 + er * m[9 - dd];
			}
			c.R1[x] = a2;
		}
		c.be[ed_idx(c.cut, z)] = make_double2(c.bnext[z], e);
	}
}

// The speculative walk of segment t of row j: pixels [zs, ze) in scan order,
// started `warm` pixels early with no in-row carry.  The pixels of the
// segment keep their errors and levels; the states at zs and after the last
// pixel go to LDS for the checks.  Position (r, q) of pixel z in the
// transposed row advances without divisions.
template <int TABM, bool HAS_M1, class Lds>
__device__ __forceinline__ void ed_walk_seg(const EdCtx &c, Lds &S, unsigned short *lv, int t, int warm)
{
	const int W = c.W, seg = c.cut.seg, nsp = c.cut.nsp;
	const int zs = t * seg, ze = min(W, zs + seg);
	double p1 = 0.0, p2 = 0.0;
	int level;
	// Warm-up: pixels zs - warm .. zs - 1.  Every lane starts at the same
	// place within a segment (r0), so the wrap to the next segment comes at
	// the same step for all of them (a uniform test) and the transposed
	// address moves by nsp, or back to the next segment's first pixel.
	// Pixels before the row's start (the first lanes) read the guard
	// element, whose error is 0.0: the state reaches pixel 0 at zero, as
	// the row starts with no in-row carry.
	const int r0 = (seg - warm % seg) % seg;
	int a = r0 * nsp + (zs - warm - r0) / seg;  // (z0 - r0) is a multiple of seg
	const int wrap = 1 - (seg - 1) * nsp, guard = seg * nsp;
	int r = r0, z = zs - warm;
	double2 v = c.be[z < 0 ? guard : a];
#pragma unroll 8
	for (int i = 0; i < warm; i++, z++) {
		// The next pixel's address and sample, read a step ahead.
		if (++r == seg) {
			r = 0;
			a += wrap;
		} else {
			a += nsp;
		}
		const double2 nv = c.be[z + 1 < 0 ? guard : a];
		const double e = ed_step<TABM, HAS_M1>(c.d, v, p1, p2, c.m0, c.m1, level);
		// Before the row's start the guard already gives zeros; the walk
		// with register records also sets them explicitly, a form that
		// runs about 1.8x faster there on gfx950 (measured on 3072-wide
		// rows), the LDS-record walks run faster without.
		const bool zero = TABM == 2 && z < 0;
		p2 = zero ? 0.0 : p1;
		p1 = zero ? 0.0 : e;
		v = nv;
	}
	S.sw[t][0] = p1;
	S.sw[t][1] = p2;
	// The segment's own pixels: q == t, r = 0 .. ze - zs - 1 (v holds the
	// first, at address t: where the warm-up's addresses end).
	const int n = ze - zs;
	for (int k = 0; k < n; k++) {
		const double2 nv = c.be[min(k + 1, n - 1) * nsp + t];
		const double e = ed_step<TABM, HAS_M1>(c.d, v, p1, p2, c.m0, c.m1, level);
		c.err[k * nsp + t] = e;
		lv[k * nsp + t] = (unsigned short)level;
		p2 = p1;
		p1 = e;
		v = nv;
	}
	S.se[t][0] = p1;
	S.se[t][1] = p2;
}

// Wave 0: the boundary checks, then (lane 0) the true chain through every
// segment whose start state differs from its predecessor's end state.
template <int TABM, bool HAS_M1, class Lds>
__device__ __forceinline__ void ed_resolve(const EdCtx &c, Lds &S, unsigned short *lv)
{
	const int lane = threadIdx.x;
	const int nseg = c.cut.nseg, seg = c.cut.seg, nsp = c.cut.nsp;
	for (int k0 = 0; k0 < nseg; k0 += 64) {
		const int k = k0 + lane;
		const bool bad = k >= 1 && k < nseg &&
			!(S.se[k - 1][0] == S.sw[k][0] && (!HAS_M1 || S.se[k - 1][1] == S.sw[k][1]));
		const unsigned long long mk = __ballot(bad);
		if (lane == 0) S.fail[k0 >> 6] = mk;
	}
	if (lane != 0) return;
	auto next_fail = [&](int k) {
		while (k < nseg) {
			const unsigned long long mk = S.fail[k >> 6] >> (k & 63);
			if (mk) return k + __ffsll((unsigned long long)mk) - 1;
			k = (k | 63) + 1;
		}
		return nseg;
	};
	const int W = c.W;
	int k = next_fail(1);
	while (k < nseg) {
		// Segments before k are verified, so segment k-1's end state is the
		// true chain's.  Walk on until the state equals the speculative one
		// at the same pixel (for the x+2 carry, both errors, from the same
		// segment's walk); that segment is right from there to its end.
		double p1 = S.se[k - 1][0], p2 = S.se[k - 1][1];
		double pe = 0.0, po = 0.0;
		int q = k, r = 0;  // pixel z = q * seg + r
		bool met = false;
		for (int z = k * seg; z < W; z++) {
			const int ix = r * nsp + q;
			int level;
			const double old = c.err[ix];
			const double e = ed_step<TABM, HAS_M1>(c.d, c.be[ix], p1, p2, c.m0, c.m1, level);
			c.err[ix] = e;
			lv[ix] = (unsigned short)level;
			met = e == old && (!HAS_M1 || (r > 0 && pe == po));
			pe = e;
			po = old;
			p2 = p1;
			p1 = e;
			if (met) break;
			if (++r == seg) {
				r = 0;
				q++;
			}
		}
		if (!met) break;
		k = next_fail(q + 1);
	}
}

// One workgroup per plane: planes[b] = (plan, output channel).  Per row:
// waves 0-3 walk the segments while waves 4-7 store the previous row's codes
// and load the next row's samples; wave 0 checks the boundaries and walks
// what differs; then every thread runs stage A for the next row.
template <int TABM, bool HAS_M1, bool ROWS>
__global__ void __launch_bounds__(ED_THREADS) errdiff_kernel(const DevPlan *__restrict__ plans,
	const int2 *__restrict__ planes)
{
	__shared__ EdLds<TABM, ROWS> S;
	const int tid = threadIdx.x;
	const int2 pl = planes[blockIdx.x];
	const DevPlan &P = plans[pl.x];
	const int o = pl.y;
	const DevOChan &q = P.oc[o];
	const double *m = c_ed_matrix[(q.subtype >= 0 && q.subtype <= 7) ? q.subtype : 0];
	const int W = P.out_w, H = P.out_h, Pq = q.P;

	EdCtx c;
	c.W = W;
	c.o = o;
	c.m0 = m[0];
	c.m1 = m[1];
	c.cut = ed_cut(W, P.ed_seg);
	c.samp = (g_cf64 *)(P.ed_samples + (size_t)q.ed_index * H * W);
	c.code = (g_cf64 *)q.code;
	c.out = (g_u8 *)P.out;
	c.out_bpr = P.out_bpr;
	c.out_nch = P.out_nch;
	c.out_depth = P.out_depth;
	if (ROWS) {
		c.R1 = S.R1;
		c.bnext = S.bnext;
		c.err = S.err;
		c.be = S.be;
		c.lvl0 = c.lvl1 = nullptr;
	} else {
		const EdRowsOff ro = ed_rows_off(W, P.ed_seg);
		double *rows = P.ed_err + (size_t)q.ed_index * ro.total;
		c.R1 = rows;
		c.bnext = rows + ro.bnext;
		c.err = rows + ro.err;
		c.be = (double2 *)(rows + ro.be);
		c.lvl0 = (unsigned short *)(rows + ro.lvl0);
		c.lvl1 = (unsigned short *)(rows + ro.lvl1);
	}
	c.d.P = Pq;
	c.d.T = q.T;
	c.d.U = q.U;
	c.d.M = q.M;
	c.d.lin = q.lin;
	if (TABM == 2)
		ed_small_fill(c.d.sm, Pq, q.T, q.U, q.M, q.lin);
	if (TABM == 1) {
		for (int k = tid; k <= Pq + 2; k += ED_THREADS) {
			EdRec r;
			if (k <= Pq) {
				r.T = k < Pq ? ((g_cf64 *)q.T)[k] : 2.0;
				r.U = ((g_cf64 *)q.U)[k];
				r.M = ((g_cf64 *)q.M)[k];
				r.lin = ((g_cf64 *)q.lin)[k];
			} else {
				r.T = 2.0;
				r.U = 2.0;
				r.M = 2.0;
				r.lin = ((g_cf64 *)q.lin)[Pq];
			}
			S.rec[k] = r;
		}
		c.d.rec = S.rec;
	} else {
		c.d.rec = nullptr;
	}
	if (TABM != 2)
		for (int k = tid; k <= ED_NB; k += ED_THREADS)
			S.Q[k] = (unsigned short)level_in(q.T, 0, Pq, k / (double)ED_NB);
	c.d.Q = S.Q;
	for (int x = tid; x < W; x += ED_THREADS) c.R1[x] = 0.0;
	if (tid == 0) c.be[c.cut.seg * c.cut.nsp] = make_double2(0.0, 0.0);  // the guard (nothing else writes it)
	ed_load_row(c, 0, tid, ED_THREADS);
	__syncthreads();
	ed_stage_a(c, m, 0);
	__syncthreads();

	// The level row of row j (two alternate), indexed straight from the LDS
	// array so that its accesses stay LDS instructions.
	auto lvl_of = [&](int jj) -> unsigned short * {
		if constexpr (ROWS) return &S.lvl[jj & 1][0];
		else return (jj & 1) ? c.lvl1 : c.lvl0;
	};
	const int warm = q.ed_warm;
	for (int j = 0; j < H; j++) {
		if (tid < c.cut.nseg) {
			ed_walk_seg<TABM, HAS_M1>(c, S, lvl_of(j), tid, warm);
		} else if (tid >= ED_WALKERS) {
			if (j > 0) ed_store_row(c, lvl_of(j - 1), j - 1, tid - ED_WALKERS, ED_THREADS - ED_WALKERS);
			if (j + 1 < H) ed_load_row(c, j + 1, tid - ED_WALKERS, ED_THREADS - ED_WALKERS);
		}
		__syncthreads();
		if (tid < 64) ed_resolve<TABM, HAS_M1>(c, S, lvl_of(j));
		__syncthreads();
		if (j + 1 < H) {
			ed_stage_a(c, m, j + 1);
			__syncthreads();
		}
	}
	ed_store_row(c, lvl_of(H - 1), H - 1, tid, ED_THREADS);
	__syncthreads();
	if (P.ed_done) {
		// Completion of the request: the last of its planes to finish copies
		// the whole output to the host and raises the flag.
		__shared__ int last;
		if (tid == 0) {
			__threadfence();
			last = atomicAdd(P.ed_done, 1) == P.n_ed - 1;
			__threadfence();
		}
		__syncthreads();
		if (last) {
			const size_t n16 = P.out_bytes / 16;
			const uint4 *src = (const uint4 *)P.out;
			uint4 *dst = (uint4 *)P.host_out;
			for (size_t k = tid; k < n16; k += ED_THREADS) dst[k] = src[k];
			for (size_t k = n16 * 16 + tid; k < P.out_bytes; k += ED_THREADS) P.host_out[k] = P.out[k];
			__threadfence_system();
			__syncthreads();
			if (tid == 0) {
				__threadfence_system();
				__hip_atomic_store(P.host_flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
			}
		}
	}
}

// ---- device memory and stream -------------------------------------------

struct Runtime {
	std::mutex mu;
	bool init = false;
	bool init_failed = false;
	int ndev = 0;
	// Free device blocks by size, reused across jobs: hipMalloc/hipFree cost
	// far more than a small image's kernels.
	std::multimap<size_t, void *> free_blocks;
	size_t cached_bytes = 0;
	size_t band_lds_max = 64 * 1024;  // dynamic LDS a band kernel may use
};

Runtime &rt()
{
	static Runtime r;
	return r;
}

constexpr size_t kCacheLimit = (size_t)8 << 30;

bool runtime_init(iw_context *ctx)
{
	Runtime &R = rt();
	std::lock_guard<std::mutex> lock(R.mu);
	if (R.init) return true;
	if (R.init_failed || hipGetDeviceCount(&R.ndev) != hipSuccess || R.ndev < 1) {
		R.init_failed = true;
		iw_set_error(ctx, "No HIP device available for the pixel pipeline");
		return false;
	}
	if (iwx::prng_mont_minv() != kMontMinv) {
		iw_set_error(ctx, "Internal: PRNG Montgomery constant mismatch");
		return false;
	}
	// The band kernels size their LDS per launch, up to the workgroup limit
	// (160 KB on gfx950); where the runtime refuses, they keep 64 KB.
	const void *bk[3] = {(const void *)band_kernel<1>, (const void *)band_kernel<2>, (const void *)band_kernel<4>};
	R.band_lds_max = 160 * 1024;
	for (const void *f : bk) {
		if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)R.band_lds_max) != hipSuccess) {
			(void)hipGetLastError();
			R.band_lds_max = 64 * 1024;
		}
	}
	R.init = true;
	return true;
}

void *dev_alloc(size_t bytes)
{
	Runtime &R = rt();
	{
		std::lock_guard<std::mutex> lock(R.mu);
		auto it = R.free_blocks.lower_bound(bytes);
		if (it != R.free_blocks.end() && it->first <= bytes + bytes / 2 + (1 << 20)) {
			void *p = it->second;
			R.cached_bytes -= it->first;
			R.free_blocks.erase(it);
			return p;
		}
	}
	void *p = nullptr;
	if (hipMalloc(&p, bytes) != hipSuccess) {
		// Release the cache and retry once.  hipFree waits for the device,
		// so the blocks are taken out under the lock and freed outside it.
		std::multimap<size_t, void *> blocks;
		{
			std::lock_guard<std::mutex> lock(R.mu);
			blocks.swap(R.free_blocks);
			R.cached_bytes = 0;
		}
		for (auto &b : blocks) hipFree(b.second);
		if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
	}
	return p;
}

void dev_release(void *p, size_t bytes)
{
	if (!p) return;
	Runtime &R = rt();
	{
		std::lock_guard<std::mutex> lock(R.mu);
		if (R.cached_bytes + bytes <= kCacheLimit) {
			R.free_blocks.emplace(bytes, p);
			R.cached_bytes += bytes;
			return;
		}
	}
	hipFree(p);
}

// Bump layout inside one device block.
struct Layout {
	size_t off = 0;
	template <class T>
	size_t add(size_t count)
	{
		off = (off + 255) & ~(size_t)255;
		size_t at = off;
		off += count * sizeof(T);
		return at;
	}
};

// ---- output optimiser scans (src/imagew-opt.c) ------------------------------
// One pass over each finished 8/16-bit output image before it is copied
// back.  It clears the colour of fully transparent pixels in place
// (:1112-1163) and forms what the optimiser decides by: colour,
// transparency, partial transparency and 16-bit precision (:18-241, :358-389); for
// each form a binary-transparency image can take (RGB or gray, 8 or 16
// bits, :418-683), the values of its varying colour-key component among
// opaque pixels; and the set of distinct 8-bit colours, as the palette
// search sees them after any 16->8 reduction, while there are at most 256
// (:719-797).  The host optimiser then only builds the new image.
//
// A workgroup takes one chunk of OPT_CHUNK pixels of one image: flags and
// key sets in LDS words, colours in an LDS hash set, then one merge into the
// image's result with global atomics.  HBM-bound: each pixel is read once
// and written only where a transparent pixel's colour is cleared.

#define OPT_WANT_KEYS 2u  // = IWX_SCAN_HAVE_KEYS
#define OPT_WANT_PAL 4u   // = IWX_SCAN_HAVE_PAL
#define OPT_THREADS 256
#define OPT_PPT 16
#define OPT_CHUNK (OPT_THREADS * OPT_PPT)
#define OPT_LSLOTS 512

static_assert(sizeof(iwx_opt_scan) % 16 == 0, "result records are cleared in 16-byte words");

struct OptJob {
	unsigned char *px;        // width*height pixels of nc samples, rows unpadded,
	                          // 16-bit samples big-endian
	iwx_opt_scan *res;
	unsigned long long npx;
	unsigned int chunk0;      // first chunk of this image in the launch
	int nc, b16;
	unsigned int want;        // IWX_SCAN_HAVE_* beyond the flags
};

struct OptShared {
	unsigned long long tab[OPT_LSLOTS];
	unsigned int used[4][8];
	unsigned int fl, pal_n, pal_over;
	int job;
};

__device__ __forceinline__ unsigned int opt_hash(unsigned int c, unsigned int slots)
{
	return (c * 2654435761u) >> (32 - __builtin_ctz(slots));
}

// Enters colour c into an open-addressed set of `slots` entries: 1 when it
// is new, 0 when present, -1 when the set is full.
__device__ __forceinline__ int opt_set_add(unsigned long long *tab, unsigned int slots, unsigned int c)
{
	const unsigned long long key = (1ull << 32) | c;
	unsigned int h = opt_hash(c, slots);
	for (unsigned int n = 0; n < slots; n++, h = (h + 1) & (slots - 1)) {
		unsigned long long s = tab[h];
		if (s == 0) {
			s = atomicCAS(&tab[h], 0ull, key);
			if (s == 0) return 1;
		}
		if (s == key) return 0;
	}
	return -1;
}

__device__ __forceinline__ void opt_bit(unsigned int *set, unsigned int v)
{
	const unsigned int b = 1u << (v & 31);
	unsigned int *w = set + (v >> 5);
	if (!(*(volatile unsigned int *)w & b)) atomicOr(w, b);
}

// One pixel's bytes as they lie in memory (little-endian word), and its
// samples (16-bit ones stored big-endian).
template <int NC, int B16>
using OptRaw = typename std::conditional<(NC * (B16 ? 2 : 1) > 4), unsigned long long, unsigned int>::type;

template <int NC, int B16>
__device__ __forceinline__ OptRaw<NC, B16> opt_raw(const g_u8 *q)
{
	constexpr int PB = NC * (B16 ? 2 : 1);
	typedef const __attribute__((address_space(1))) unsigned int g_cu32;
	typedef const __attribute__((address_space(1))) unsigned long long g_cu64;
	if constexpr (PB == 8) return *(g_cu64 *)q;
	else if constexpr (PB == 4) return *(g_cu32 *)q;
	else if constexpr (PB == 2) return *(g_cu16 *)q;
	else {
		OptRaw<NC, B16> w = 0;
#pragma unroll
		for (int k = 0; k < PB; k++) w |= (OptRaw<NC, B16>)q[k] << (8 * k);
		return w;
	}
}

template <int NC, int B16>
__device__ __forceinline__ void opt_decode(OptRaw<NC, B16> w, unsigned int s[4])
{
#pragma unroll
	for (int k = 0; k < NC; k++) {
		if constexpr (B16) s[k] = (unsigned int)(((w >> (16 * k)) & 0xffu) << 8) | (unsigned int)((w >> (16 * k + 8)) & 0xffu);
		else s[k] = (unsigned int)((w >> (8 * k)) & 0xffu);
	}
}

template <int NC, int B16>
__device__ __forceinline__ void opt_clear_px(g_u8 *q)
{
	constexpr int PB = NC * (B16 ? 2 : 1);
	if constexpr (PB == 4) *(__attribute__((address_space(1))) unsigned int *)q = 0u;
	else if constexpr (PB == 8) *(__attribute__((address_space(1))) unsigned long long *)q = 0ull;
	else if constexpr (PB == 2) *(__attribute__((address_space(1))) unsigned short *)q = 0;
	else {
#pragma unroll
		for (int k = 0; k < PB; k++) q[k] = 0;
	}
}

template <int NC, int B16>
__device__ void opt_scan_chunk(const OptJob &J, unsigned long long first, OptShared &S)
{
	constexpr int PB = NC * (B16 ? 2 : 1);
	constexpr bool ALPHA = (NC == 2 || NC == 4), RGB = (NC >= 3);
	constexpr unsigned int AMAX = B16 ? 65535u : 255u;
	const bool keys = ALPHA && (J.want & OPT_WANT_KEYS);
	const bool pal = (J.want & OPT_WANT_PAL) != 0;
	unsigned int fl = 0;
	// Loads of OPT_BATCH pixels are issued together, then worked through.
	constexpr int OPT_BATCH = 8;
	for (int k0 = 0; k0 < OPT_PPT; k0 += OPT_BATCH) {
	OptRaw<NC, B16> raw[OPT_BATCH];
#pragma unroll
	for (int k = 0; k < OPT_BATCH; k++) {
		const unsigned long long p = first + (unsigned long long)(k0 + k) * OPT_THREADS + threadIdx.x;
		raw[k] = p < J.npx ? opt_raw<NC, B16>((const g_u8 *)J.px + p * PB) : 0;
	}
#pragma unroll
	for (int k = 0; k < OPT_BATCH; k++) {
		const unsigned long long p = first + (unsigned long long)(k0 + k) * OPT_THREADS + threadIdx.x;
		if (p >= J.npx) break;
		g_u8 *q = (g_u8 *)J.px + p * PB;
		unsigned int s[4];
		opt_decode<NC, B16>(raw[k], s);
		const unsigned int a = ALPHA ? s[NC - 1] : AMAX;
		if (ALPHA && a == 0) {
			bool any = false;
#pragma unroll
			for (int c = 0; c < NC - 1; c++) any |= s[c] != 0;
			if (any) {
				opt_clear_px<NC, B16>(q);
#pragma unroll
				for (int c = 0; c < NC - 1; c++) s[c] = 0;
			}
		}
		if (ALPHA && a < AMAX) fl |= IWX_SCAN_TRNS;
		if (ALPHA && a > 0 && a < AMAX) fl |= IWX_SCAN_PARTIAL;
		if (B16) {
#pragma unroll
			for (int c = 0; c < NC; c++)
				if ((s[c] >> 8) != (s[c] & 0xffu)) fl |= IWX_SCAN_16BIT;
		}
		if (RGB && (s[0] != s[1] || s[0] != s[2])) fl |= IWX_SCAN_COLOR;
		const unsigned int r8 = B16 ? s[0] >> 8 : s[0];
		const unsigned int g8 = RGB ? (B16 ? s[1] >> 8 : s[1]) : r8;
		const unsigned int b8 = RGB ? (B16 ? s[2] >> 8 : s[2]) : r8;
		if (keys && a != 0) {
			opt_bit(S.used[IWX_KEY_GRAY8], r8);
			if (RGB && g8 == 192 && b8 == 192) opt_bit(S.used[IWX_KEY_RGB8], r8);
			if (B16 && r8 == 192) {
				opt_bit(S.used[IWX_KEY_GRAY16], s[0] & 0xffu);
				if (RGB && s[1] == 192u * 257u && s[2] == 192u * 257u) opt_bit(S.used[IWX_KEY_RGB16], s[0] & 0xffu);
			}
		}
		if (pal && !*(volatile unsigned int *)&S.pal_over) {
			const unsigned int a8 = ALPHA ? (B16 ? a >> 8 : a) : 255u;
			const unsigned int c = a8 == 0 ? 0u : r8 | (g8 << 8) | (b8 << 16) | (a8 << 24);
			const int r = opt_set_add(S.tab, OPT_LSLOTS, c);
			if (r < 0 || (r > 0 && atomicAdd(&S.pal_n, 1u) >= 256u)) *(volatile unsigned int *)&S.pal_over = 1u;
		}
	}
	}
	if (fl) atomicOr(&S.fl, fl);
}

__global__ __launch_bounds__(OPT_THREADS) void opt_scan_kernel(const OptJob *__restrict__ jobs, int njobs)
{
	__shared__ OptShared S;
	const unsigned int chunk = blockIdx.x;
	if (threadIdx.x == 0) {
		// The last image whose first chunk is at or before this one.
		int lo = 0, hi = njobs - 1;
		while (lo < hi) {
			const int mid = (lo + hi + 1) >> 1;
			if (jobs[mid].chunk0 <= chunk) lo = mid;
			else hi = mid - 1;
		}
		S.job = lo;
		S.fl = 0;
		S.pal_n = 0;
		S.pal_over = __hip_atomic_load(&jobs[lo].res->pal_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
	for (int i = threadIdx.x; i < OPT_LSLOTS; i += OPT_THREADS) S.tab[i] = 0;
	if (threadIdx.x < 32) (&S.used[0][0])[threadIdx.x] = 0;
	__syncthreads();
	const OptJob J = jobs[S.job];
	const unsigned long long first = (unsigned long long)(chunk - J.chunk0) * OPT_CHUNK;
	switch (J.nc * 2 + J.b16) {
	case 2: opt_scan_chunk<1, 0>(J, first, S); break;
	case 3: opt_scan_chunk<1, 1>(J, first, S); break;
	case 4: opt_scan_chunk<2, 0>(J, first, S); break;
	case 5: opt_scan_chunk<2, 1>(J, first, S); break;
	case 6: opt_scan_chunk<3, 0>(J, first, S); break;
	case 7: opt_scan_chunk<3, 1>(J, first, S); break;
	case 8: opt_scan_chunk<4, 0>(J, first, S); break;
	case 9: opt_scan_chunk<4, 1>(J, first, S); break;
	default: break;
	}
	__syncthreads();
	iwx_opt_scan *R = J.res;
	if (threadIdx.x < 32) {
		const unsigned int u = (&S.used[0][0])[threadIdx.x];
		if (u) atomicOr(&(&R->used[0][0])[threadIdx.x], u);
	}
	if (threadIdx.x == 0 && S.fl) atomicOr(&R->flags, S.fl);
	if (!(J.want & OPT_WANT_PAL)) return;
	if (S.pal_over) {
		if (threadIdx.x == 0) atomicOr(&R->pal_over, 1u);
		return;
	}
	for (int i = threadIdx.x; i < OPT_LSLOTS; i += OPT_THREADS) {
		const unsigned long long v = S.tab[i];
		if (!v) continue;
		if (__hip_atomic_load(&R->pal_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
		const int r = opt_set_add(R->pal, IWX_SCAN_PAL_SLOTS, (unsigned int)v);
		if (r < 0 || (r > 0 && atomicAdd(&R->pal_count, 1u) >= 256u)) {
			atomicOr(&R->pal_over, 1u);
			break;
		}
	}
}

__global__ void opt_clear_kernel(uint4 *p, size_t n)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Which of the optimiser's inputs the device forms for this prepared
// context (IWX_SCAN_HAVE_*); 0: the host scans.  Negated output is scanned
// by the host after the negation (src/imagew-main.c:1439-1440).
unsigned int opt_scan_want(const iw_context *ctx)
{
	const iw_image &im = ctx->img2;
	if (im.sampletype != IW_SAMPLETYPE_UINT || ctx->reduced_output_maxcolor_flag || ctx->req.negate_target)
		return 0;
	if (im.bit_depth != 8 && im.bit_depth != 16) return 0;
	const int t = im.imgtype;
	if (t != IW_IMGTYPE_GRAY && t != IW_IMGTYPE_GRAYA && t != IW_IMGTYPE_RGB && t != IW_IMGTYPE_RGBA) return 0;
	if (im.bpr != (size_t)im.width * iw_imgtype_num_channels(t) * (im.bit_depth / 8)) return 0;
	if ((size_t)im.width * im.height == 0) return 0;
	unsigned int w = IWX_SCAN_HAVE_FLAGS;
	if (IW_IMGTYPE_HAS_ALPHA(t) && (ctx->output_profile & IW_PROFILE_BINARYTRNS) && ctx->opt_binary_trns)
		w |= IWX_SCAN_HAVE_KEYS;
	if (ctx->output_profile & (IW_PROFILE_PAL1 | IW_PROFILE_PAL2 | IW_PROFILE_PAL4 | IW_PROFILE_PAL8 |
			IW_PROFILE_GRAY1 | IW_PROFILE_GRAY2 | IW_PROFILE_GRAY4))
		w |= IWX_SCAN_HAVE_PAL;
	return w;
}

OptJob opt_job(const iw_context *ctx, void *px, iwx_opt_scan *res, unsigned int want)
{
	OptJob J;
	J.px = (unsigned char *)px;
	J.res = res;
	J.npx = (unsigned long long)ctx->img2.width * ctx->img2.height;
	J.chunk0 = 0;
	J.nc = iw_imgtype_num_channels(ctx->img2.imgtype);
	J.b16 = ctx->img2.bit_depth == 16;
	J.want = want;
	return J;
}

// Bytes of device memory for the scans of n images: descriptors, then one
// result per image.
size_t opt_scan_bytes(size_t n)
{
	return ((n * sizeof(OptJob) + 255) & ~(size_t)255) + n * sizeof(iwx_opt_scan);
}

// Clears n results at `region` + descriptors' size and launches the scans
// of `jobs` (each one's res in that region) on st.  `jobs` must live until
// the stream has passed the launch.
bool opt_scan_enqueue(std::vector<OptJob> &jobs, void *region, size_t n, hipStream_t st)
{
	OptJob *dj = (OptJob *)region;
	iwx_opt_scan *res = (iwx_opt_scan *)((char *)region + ((n * sizeof(OptJob) + 255) & ~(size_t)255));
	unsigned long long chunks = 0;
	for (auto &J : jobs) {
		J.chunk0 = (unsigned int)chunks;
		chunks += (J.npx + OPT_CHUNK - 1) / OPT_CHUNK;
		if (J.res < res || J.res >= res + n) return false;
	}
	if (chunks >= 0x7fffffffull) return false;
	const size_t n16 = n * sizeof(iwx_opt_scan) / 16;
	if (n16) {
		const unsigned int g = (unsigned int)std::min<size_t>(1024, (n16 + 255) / 256);
		opt_clear_kernel<<<g, 256, 0, st>>>((uint4 *)res, n16);
	}
	if (!jobs.empty() && chunks) {
		if (hipMemcpyAsync(dj, jobs.data(), jobs.size() * sizeof(OptJob), hipMemcpyHostToDevice, st) != hipSuccess)
			return false;
		opt_scan_kernel<<<(unsigned int)chunks, OPT_THREADS, 0, st>>>(dj, (int)jobs.size());
	}
	return hipGetLastError() == hipSuccess;
}

}  // namespace

// ---- a device-resident job --------------------------------------------------

struct iwamd_job {
	iw_context *ctx = nullptr;
	iwamd_plan plan;
	DevPlan D;
	void *block = nullptr;
	size_t block_bytes = 0;
	size_t out_off = 0, out_bytes = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
	int n_ed = 0;
	int ed_list[4] = {-1, -1, -1, -1};
	double out_pixels = 0.0, alg_bytes = 0.0, flops = 0.0;
	int band_rows = 0;  // rows per band of the fused kernel; 0: two-pass kernels
	bool ran = false;
	bool skip_pixels = false;  // unknown filter: the reference reports and outputs zeros
	void *ws = nullptr;        // the job's work set (WorkSet)
};

namespace {

// Host side of one job's device data: the tables, where each lives in the
// job's part of a device block, and the resampling path.
struct WOff { size_t start, src, w, copy, dfirst, dcount, doff, dw; };
struct QOff { size_t T, U, lin, code, nc, ix, D, prow, pcol, M; };

struct Staged {
	std::vector<iwx::WeightList> wv, wh;
	std::vector<iwx::DenseWeights> dv, dh;
	std::vector<FastItem> fast;  // chosen-precision tiles (job index set when placed); empty: plain
	std::vector<std::vector<double>> luts;
	int ic_lut[4][3];
	int alpha_lut = -1;
	std::vector<iwx::QuantTable> qt;
	std::vector<iwx::EdTables> et;
	std::vector<iwx::PrngTables> pt;
	int ed_of[4] = {-1, -1, -1, -1};
	int n_ed = 0;
	int ed_warm = 0;           // error diffusion: warm-up pixels (0: by the channel's weights and levels)
	int ed_seg = ED_SEG_MIN;   // error diffusion: fewest pixels per segment
	int ed_list[4] = {-1, -1, -1, -1};
	bool in_int = true;
	bool in_elsewhere = false; // the input image is already in device memory (no region, no copy)
	std::vector<double> in_f64;  // 32-bit float input in linear light (non-linear colour spaces), else empty
	size_t in_f64_bpr = 0;
	bool skip_pixels = false;  // unknown filter: the reference reports and outputs zeros
	int band_rows = 0;         // rows per band of the fused kernel; 0: two-pass kernels
	int fast_ti = 0;           // output columns per chosen-precision tile
	std::vector<int4> h4i;     // chosen form: the horizontal list by 4-column blocks (DevPlan::h4i)
	std::vector<float> h4w;
	size_t o_h4i = 0, o_h4w = 0;
	bool shape_forced = false; // band height or tile width set by an option (not tuned)
	size_t in_bytes = 0, out_bytes = 0;
	size_t o_in = 0, o_inter = 0, o_out = 0, o_eds = 0, o_ede = 0;
	std::vector<size_t> o_lut;
	std::vector<WOff> ov, oh;
	std::vector<QOff> oq;
	size_t bytes = 0;          // size of the job's part of a block
	size_t o_tables = 0;       // start of the host-built tables (through the end of the part)
	std::vector<uint8_t> pack;  // their host image, sent in one copy (lives until the stream syncs)
};

size_t band_lds_bytes(int R, int C, int W)
{
	return sizeof(double) * (size_t)R * BAND_MAX_K + sizeof(float) * (size_t)C * R * W;
}

// Whether bands of R output rows fit the band kernel: LDS for the rows and
// input spans of at most BAND_MAX_K positions.
bool band_fits(const iwamd_plan &p, const std::vector<iwx::DenseWeights> &dv, int R)
{
	if (band_lds_bytes(R, p.inter_nch, p.inter_width) > rt().band_lds_max) return false;
	for (int c = 0; c < p.inter_nch; c++) {
		const iwx::DenseWeights &d = dv[c];
		for (int j0 = 0; j0 < p.inter_height; j0 += R) {
			int lo = 0x7fffffff, hi = -0x7fffffff;
			for (int r = 0; r < R && j0 + r < p.inter_height; r++) {
				if (!d.count[j0 + r]) continue;
				lo = std::min(lo, d.first[j0 + r]);
				hi = std::max(hi, d.first[j0 + r] + d.count[j0 + r] - 1);
			}
			if (hi >= lo && hi - lo + 1 > BAND_MAX_K) return false;
		}
	}
	return true;
}

// The tallest band that fits (0: none).
int choose_band_rows(const iwamd_plan &p, const std::vector<iwx::DenseWeights> &dv)
{
	for (int R : {4, 2, 1})
		if (band_fits(p, dv, R)) return R;
	return 0;
}

bool same_dense(const iwx::DenseWeights &a, const iwx::DenseWeights &b)
{
	return a.first == b.first && a.count == b.count && a.off == b.off && a.w == b.w;
}

// Positions [lo, lo + cnt) that rows a..a+n-1 of d reach (cnt 0: none).
void dense_span(const iwx::DenseWeights &d, int a, int n, int &lo, int &cnt)
{
	int l = 0x7fffffff, h = -0x7fffffff;
	for (int k = a; k < a + n; k++) {
		if (!d.count[k]) continue;
		l = std::min(l, d.first[k]);
		h = std::max(h, d.first[k] + d.count[k] - 1);
	}
	if (h < l) {
		l = 0;
		h = -1;
	}
	lo = l;
	cnt = h - l + 1;
}

// Intermediate columns a tile reads: the whole 16-column groups (absolute,
// so the same for both tile widths) that its lists reach.
void fast_tile_span(const iwx::DenseWeights &dh, int i0, int ni, int &lo, int &cnt)
{
	dense_span(dh, i0, ni, lo, cnt);
	if (cnt > 0) {
		const int hi = lo + cnt - 1;
		lo &= ~15;
		cnt = (hi & ~15) + 16 - lo;
	}
}

// Tiles of the chosen-precision form, or none where it does not apply: it
// covers integer outputs (8 or 16 bits) whose channels share one weight list
// per dimension, with bands of at most FAST_MAX_KV input rows.  want_ti: tile
// width (32 or 16 output columns; 0: 32).
void plan_fast(const iwamd_plan &p, Staged &S, int want_ti = 0)
{
	S.fast.clear();
	if (p.out_sampletype == IW_SAMPLETYPE_FLOATINGPOINT || p.out_bit_depth > 16) return;
	// Floating-point inputs can hold values past the form's range.
	if (p.in_sampletype == IW_SAMPLETYPE_FLOATINGPOINT) return;
	for (int c = 1; c < p.inter_nch; c++) {
		if (!same_dense(S.dv[c], S.dv[0]) || !same_dense(S.dh[c], S.dh[0])) return;
		if (S.wv[c].edge_value != S.wv[0].edge_value) return;
	}
	const iwx::DenseWeights &dv = S.dv[0], &dh = S.dh[0];
	for (int j0 = 0; j0 < p.inter_height; j0 += FAST_ROWS) {
		int lo, n;
		dense_span(dv, j0, std::min(FAST_ROWS, p.inter_height - j0), lo, n);
		if (n > FAST_MAX_KV) return;
	}
	const int ti = want_ti == 16 ? 16 : 32;
	S.fast_ti = ti;
	for (int j0 = 0; j0 < p.inter_height; j0 += FAST_ROWS) {
		FastItem it;
		it.job = 0;
		it.j0 = j0;
		dense_span(dv, j0, std::min(FAST_ROWS, p.inter_height - j0), it.p_lo, it.kv);
		for (int i0 = 0; i0 < p.out_width; i0 += ti) {
			it.i0 = i0;
			it.ni = std::min(ti, p.out_width - i0);
			fast_tile_span(dh, i0, it.ni, it.x_lo, it.nx);
			S.fast.push_back(it);
		}
	}
}

// The shared horizontal list by 4-column blocks (DevPlan::h4i, h4w).
void build_h4(const iwx::DenseWeights &dh, Staged &S)
{
	const int n = (int)dh.first.size();
	S.h4i.assign(n, make_int4(0, 0, 0, 0));
	S.h4w.assign(4, 0.0f);  // block 0: zeros
	for (int i = 0; i < n; i++) {
		if (!dh.count[i]) continue;
		const int f = dh.first[i], e = f + dh.count[i] - 1;
		const int b0 = f >> 2, b1 = e >> 2;  // floor, also for virtual columns below 0
		S.h4i[i] = make_int4(b0, b1 - b0 + 1, (int)(S.h4w.size() / 4), 0);
		for (int b = b0; b <= b1; b++)
			for (int t = 0; t < 4; t++) {
				const int x = 4 * b + t;
				S.h4w.push_back(x >= f && x <= e ? (float)dh.w[dh.off[i] + (x - f)] : 0.0f);
			}
	}
}

// 32-bit float input in linear light as doubles (Staged::in_f64): stored
// channel k of every pixel through the transfer function of the intermediate
// channel that reads it (src/imagew-main.c:355-364; channels no intermediate
// channel converts -- alpha, or unread -- as stored).
bool build_f64_input(iw_context *ctx, const iwamd_plan &p, Staged &S, bool in_elsewhere)
{
	if (in_elsewhere) {
		iw_set_error(ctx, "Device-resident floating-point input needs a linear input colorspace");
		return false;
	}
	const int nch = p.in_nch_physical;
	std::vector<const struct iw_csdescr *> cs(nch, nullptr);
	for (int c = 0; c < p.inter_nch; c++) {
		const int nk = p.ich[c].cvt_to_grayscale ? 3 : 1;
		for (int k = 0; k < nk; k++) {
			const int ch = p.ich[c].input_channel + k;
			if (ch >= nch || p.ich[c].in_cs.cstype == IW_CSTYPE_LINEAR) continue;
			if (cs[ch] && (cs[ch]->cstype != p.ich[c].in_cs.cstype || cs[ch]->gamma != p.ich[c].in_cs.gamma)) {
				iw_set_error(ctx, "Floating-point input channel read in two colorspaces");
				return false;
			}
			cs[ch] = &p.ich[c].in_cs;
		}
	}
	const bool tr = (p.orient_transform & 4) != 0;
	const size_t rows = tr ? p.in_width : p.in_height, cols = tr ? p.in_height : p.in_width;
	S.in_f64_bpr = cols * nch * 8;
	S.in_f64.resize(rows * cols * nch);
	const uint8_t *src = (const uint8_t *)p.in_pixels;
	for (size_t y = 0; y < rows; y++)
		for (size_t x = 0; x < cols; x++)
			for (int ch = 0; ch < nch; ch++) {
				float f;
				memcpy(&f, src + y * p.in_bpr + (x * nch + ch) * 4, 4);
				const double v = (double)f;
				S.in_f64[(y * cols + x) * nch + ch] = cs[ch] ? iwx_x_to_linear(v, cs[ch]) : v;
			}
	return true;
}

// Build the host tables of a plan, choose its path and lay out its device
// data (offsets from the job's base).  path: "band", "twopass" or null.
bool stage_job(iw_context *ctx, const iwamd_plan &p, Staged &S, const char *path, bool out_elsewhere = false,
	bool in_elsewhere = false)
{
	using namespace iwx;
	// Floating-point input: 32-bit samples are read as stored; the reference
	// reads any other stored depth as 0.0 (src/imagew-main.c:267-272), which
	// every transfer function maps to 0.0.  A 32-bit sample in a non-linear
	// colour space goes through libm's pow (src/imagew-main.c:53-66,
	// :355-364), which the device does not evaluate: the host converts the
	// image to linear doubles first, each stored channel with the colour
	// space of the intermediate channel that reads it, and the kernels read
	// those.
	if (p.in_sampletype == IW_SAMPLETYPE_FLOATINGPOINT && p.in_bit_depth == 32) {
		bool nonlinear = false;
		for (int c = 0; c < p.inter_nch; c++) nonlinear |= p.ich[c].in_cs.cstype != IW_CSTYPE_LINEAR;
		if (nonlinear && !build_f64_input(ctx, p, S, in_elsewhere)) return false;
	}
	S.wv.resize(p.inter_nch);
	S.wh.resize(p.inter_nch);
	S.dv.resize(p.inter_nch);
	S.dh.resize(p.inter_nch);
	for (int c = 0; c < p.inter_nch; c++) {
		S.wv[c] = build_weight_list(p.rs[IW_DIMENSION_V], p.ich[c].channeltype, p.ich[c].edge_sample_value[IW_DIMENSION_V]);
		S.wh[c] = build_weight_list(p.rs[IW_DIMENSION_H], p.ich[c].channeltype, p.ich[c].edge_sample_value[IW_DIMENSION_H]);
		if (!S.wv[c].ok || !S.wh[c].ok) {
			// The reference reports the error and carries on
			// (src/imagew-resize.c:512-515).
			iw_set_error(ctx, "Internal: Unknown resize algorithm");
			S.skip_pixels = true;
		}
		S.dv[c] = dense_weights(S.wv[c]);
		S.dh[c] = dense_weights(S.wh[c]);
	}
	S.in_int = p.in_sampletype != IW_SAMPLETYPE_FLOATINGPOINT;
	const int nvalues = S.in_int ? (1 << p.in_bit_depth) : 0;
	std::map<std::tuple<int, double, double>, int> lut_index;
	auto get_lut = [&](double maxcc, const struct iw_csdescr &cs) {
		auto key = std::make_tuple(cs.cstype, cs.cstype == IW_CSTYPE_GAMMA ? cs.gamma : 0.0, maxcc);
		auto it = lut_index.find(key);
		if (it != lut_index.end()) return it->second;
		int id = (int)S.luts.size();
		S.luts.push_back(build_input_lut(nvalues, maxcc, cs));
		lut_index[key] = id;
		return id;
	};
	for (auto &r : S.ic_lut) r[0] = r[1] = r[2] = -1;
	if (S.in_int) {
		for (int c = 0; c < p.inter_nch; c++) {
			const iwamd_ichan &ic = p.ich[c];
			const int nk = ic.cvt_to_grayscale ? 3 : 1;
			for (int k = 0; k < nk; k++) {
				const int ch = ic.input_channel + k;
				if (ch >= p.in_nch_physical) continue;
				// Fast path: input max code; slow path: the channel's own
				// (src/imagew-main.c:307-383).
				const double mcc = p.in_ch_slow[ic.input_channel] ? p.in_ch_maxcolorcode[ch] : p.in_maxcolorcode;
				S.ic_lut[c][k] = get_lut(mcc, ic.in_cs);
			}
		}
		if (p.in_has_alpha && p.in_alpha_index < p.in_nch_physical) {
			struct iw_csdescr lin;
			iw_make_linear_csdescr(&lin);
			S.alpha_lut = get_lut(p.in_ch_maxcolorcode[p.in_alpha_index], lin);
		}
	}
	S.qt.resize(p.out_nch);
	S.et.resize(p.out_nch);
	S.pt.resize(p.out_nch);
	if (p.out_sampletype != IW_SAMPLETYPE_FLOATINGPOINT) {
		for (int o = 0; o < p.out_nch; o++) {
			S.qt[o] = build_quant_table(p.och[o], p.och[o].use_nc_table != 0,
				p.och[o].ditherfamily == IW_DITHERFAMILY_NONE);
			if (p.och[o].ditherfamily == IW_DITHERFAMILY_RANDOM) {
				S.pt[o] = build_prng_tables(p.och[o].random_seed, p.out_width, p.out_height);
				if (!S.pt[o].ok) {
					iw_set_error(ctx, "Random seed not supported by the GPU path");
					return false;
				}
			}
			if (p.och[o].ditherfamily == IW_DITHERFAMILY_ERRDIFF) {
				S.et[o] = build_ed_tables(S.qt[o]);
				if (S.qt[o].P > 65535) {
					iw_set_error(ctx, "Too many output levels for error diffusion on the GPU path");
					return false;
				}
				if (p.out_width >= (1 << 22)) {
					// ed_cut's float division is exact below 2^22 pixels.
					iw_set_error(ctx, "Rows too wide for error diffusion on the GPU path");
					return false;
				}
				S.ed_of[o] = S.n_ed;
				S.ed_list[S.n_ed++] = o;
			}
		}
	}

	// The kernels read logical rows/columns [start, start+size) of the input;
	// refuse plans that would reach outside it (the reference's crop clamp can
	// leave a non-positive size, src/imagew-main.c:1949-1958).
	if (p.input_w < 1 || p.input_h < 1 || p.input_start_x < 0 || p.input_start_y < 0 ||
		p.input_start_x + p.input_w > p.in_width || p.input_start_y + p.input_h > p.in_height) {
		iw_set_error(ctx, "Image too large to process");
		return false;
	}
	// Rows of the physical buffer: a transposing orientation swaps the
	// logical width and height (src/imagew-api.c:660-684).
	const int phys_rows = (p.orient_transform & 4) ? p.in_width : p.in_height;

	// Resampling path: the fused band kernel when a band of rows fits in LDS
	// (wide rows keep more workgroups per CU in the two-pass form), else the
	// two-pass kernels with the intermediate in HBM.
	const bool force_two = path && !strcmp(path, "twopass");
	const bool force_band = path && !strcmp(path, "band");
	// Options amd:tile (16 or 32) and amd:band_rows (1, 2 or 4) fix the shape
	// where it fits; otherwise the autotuner may pick it (tune_apply).
	auto opt_int = [&](const char *name) {
		const char *v = iw_get_option(ctx, name);
		return v ? atoi(v) : 0;
	};
	const int want_ti = opt_int("amd:tile"), want_rows = opt_int("amd:band_rows");
	// Options amd:ed_warm and amd:ed_seg set the error-diffusion walks' warm-up
	// and the fewest pixels per segment; they change the time, not the bytes.
	S.ed_warm = std::max(0, opt_int("amd:ed_warm"));
	if (opt_int("amd:ed_seg") >= 1) S.ed_seg = opt_int("amd:ed_seg");
	if (p.precision == IWAMD_PRECISION_FAST && !S.skip_pixels)
		plan_fast(p, S, (want_ti == 16 || want_ti == 32) ? want_ti : 0);
	if (!S.fast.empty()) {
		S.shape_forced = S.fast_ti == want_ti;
		build_h4(S.dh[0], S);
	}
	if (S.fast.empty()) {
		S.band_rows = force_two ? 0 : choose_band_rows(p, S.dv);
		if (!force_band && S.band_rows && p.inter_width > 2048) S.band_rows = 0;
		if (S.band_rows && (want_rows == 1 || want_rows == 2 || want_rows == 4) && band_fits(p, S.dv, want_rows)) {
			S.band_rows = want_rows;
			S.shape_forced = true;
		}
	}

	Layout L;
	S.in_bytes = S.in_f64.empty() ? p.in_bpr * (size_t)phys_rows : S.in_f64.size() * sizeof(double);
	S.in_elsewhere = in_elsewhere;
	S.o_in = L.add<uint8_t>(in_elsewhere ? 0 : S.in_bytes);
	S.o_inter = L.add<float>(S.band_rows ? 0 : (size_t)p.inter_nch * p.inter_height * p.inter_width);
	S.out_bytes = p.out_bpr * (size_t)p.out_height;
	S.o_out = out_elsewhere ? 0 : L.add<uint8_t>(S.out_bytes);
	S.o_eds = L.add<double>((size_t)S.n_ed * p.out_height * p.out_width);
	S.o_ede = L.add<double>((size_t)S.n_ed * ed_rows_off(p.out_width, S.ed_seg).total);
	S.o_tables = L.add<uint8_t>(0);
	S.o_lut.resize(S.luts.size());
	for (size_t i = 0; i < S.luts.size(); i++) S.o_lut[i] = L.add<double>(S.luts[i].size());
	S.ov.resize(p.inter_nch);
	S.oh.resize(p.inter_nch);
	auto place_w = [&](const WeightList &wl, WOff &o) {
		o.start = L.add<int>(wl.start.size());
		o.src = L.add<int>(wl.src.size());
		o.w = L.add<double>(wl.w.size());
		o.copy = L.add<int>(wl.copy_src.size());
	};
	for (int c = 0; c < p.inter_nch; c++) {
		place_w(S.wv[c], S.ov[c]);
		place_w(S.wh[c], S.oh[c]);
		S.ov[c].dfirst = L.add<int>(S.dv[c].first.size());
		S.ov[c].dcount = L.add<int>(S.dv[c].count.size());
		S.ov[c].doff = L.add<int>(S.dv[c].off.size());
		S.ov[c].dw = L.add<double>(S.dv[c].w.size());
		S.oh[c].dfirst = L.add<int>(S.dh[c].first.size());
		S.oh[c].dcount = L.add<int>(S.dh[c].count.size());
		S.oh[c].doff = L.add<int>(S.dh[c].off.size());
		S.oh[c].dw = L.add<double>(S.dh[c].w.size());
	}
	S.o_h4i = L.add<int4>(S.h4i.size());
	S.o_h4w = L.add<float>(S.h4w.size());
	S.oq.resize(p.out_nch);
	// Channels whose tables are the same (the host builds each once per
	// process) share one device copy.
	auto same_tables = [&](int a, int b) {
		const iwx::QuantTable &x = S.qt[a], &y = S.qt[b];
		return x.P == y.P && x.ix_log2 == y.ix_log2 && x.code_is_level == y.code_is_level && x.T == y.T &&
			x.U == y.U && x.lin == y.lin && x.code == y.code && x.nc == y.nc && x.ix == y.ix && x.D == y.D &&
			S.pt[a].row_q == S.pt[b].row_q && S.pt[a].col_mont == S.pt[b].col_mont && S.et[a].M == S.et[b].M;
	};
	for (int o = 0; o < (int)S.qt.size(); o++) {
		QOff &q = S.oq[o];
		int same = -1;
		for (int k = 0; k < o && same < 0; k++)
			if (same_tables(k, o)) same = k;
		if (same >= 0) {
			q = S.oq[same];
			continue;
		}
		q.T = L.add<double>(S.qt[o].T.size());
		q.U = L.add<double>(S.qt[o].U.size());
		q.lin = L.add<double>(S.qt[o].lin.size());
		q.code = L.add<double>(S.qt[o].code.size());
		q.nc = L.add<double>(S.qt[o].nc.size());
		q.ix = L.add<uint16_t>(S.qt[o].ix.size());
		q.D = L.add<double>(S.qt[o].D.size());
		q.prow = L.add<uint64_t>(S.pt[o].row_q.size());
		q.pcol = L.add<uint64_t>(S.pt[o].col_mont.size());
		q.M = L.add<double>(S.et[o].M.size());
	}
	S.bytes = (L.off + 255) & ~(size_t)255;
	return true;
}

// Queue the host-to-device copies of a staged job into its part of a block:
// the input image, then every host-built table in one copy of the tables
// region from a host image of it (S.pack).  S must live until the stream has
// been synchronised.
bool upload_job(Staged &S, const iwamd_plan &p, char *base, hipStream_t st)
{
	bool ok = true;
	if (S.in_bytes && !S.in_elsewhere)
		ok &= hipMemcpyAsync(base + S.o_in, S.in_f64.empty() ? p.in_pixels : (const void *)S.in_f64.data(), S.in_bytes,
			hipMemcpyHostToDevice, st) == hipSuccess;
	const size_t t0 = S.o_tables;
	S.pack.assign(S.bytes - t0, 0);
	auto up = [&](size_t off, const void *src, size_t bytes) {
		if (bytes) memcpy(S.pack.data() + (off - t0), src, bytes);
	};
	for (size_t i = 0; i < S.luts.size(); i++) up(S.o_lut[i], S.luts[i].data(), S.luts[i].size() * 8);
	auto up_w = [&](const iwx::WeightList &wl, const WOff &o) {
		up(o.start, wl.start.data(), wl.start.size() * 4);
		up(o.src, wl.src.data(), wl.src.size() * 4);
		up(o.w, wl.w.data(), wl.w.size() * 8);
		up(o.copy, wl.copy_src.data(), wl.copy_src.size() * 4);
	};
	for (int c = 0; c < p.inter_nch; c++) {
		up_w(S.wv[c], S.ov[c]);
		up_w(S.wh[c], S.oh[c]);
		up(S.ov[c].dfirst, S.dv[c].first.data(), S.dv[c].first.size() * 4);
		up(S.ov[c].dcount, S.dv[c].count.data(), S.dv[c].count.size() * 4);
		up(S.ov[c].doff, S.dv[c].off.data(), S.dv[c].off.size() * 4);
		up(S.ov[c].dw, S.dv[c].w.data(), S.dv[c].w.size() * 8);
		up(S.oh[c].dfirst, S.dh[c].first.data(), S.dh[c].first.size() * 4);
		up(S.oh[c].dcount, S.dh[c].count.data(), S.dh[c].count.size() * 4);
		up(S.oh[c].doff, S.dh[c].off.data(), S.dh[c].off.size() * 4);
		up(S.oh[c].dw, S.dh[c].w.data(), S.dh[c].w.size() * 8);
	}
	up(S.o_h4i, S.h4i.data(), S.h4i.size() * sizeof(int4));
	up(S.o_h4w, S.h4w.data(), S.h4w.size() * sizeof(float));
	for (int o = 0; o < (int)S.qt.size(); o++) {
		const QOff &q = S.oq[o];
		up(q.T, S.qt[o].T.data(), S.qt[o].T.size() * 8);
		up(q.U, S.qt[o].U.data(), S.qt[o].U.size() * 8);
		up(q.lin, S.qt[o].lin.data(), S.qt[o].lin.size() * 8);
		up(q.code, S.qt[o].code.data(), S.qt[o].code.size() * 8);
		up(q.nc, S.qt[o].nc.data(), S.qt[o].nc.size() * 8);
		up(q.ix, S.qt[o].ix.data(), S.qt[o].ix.size() * 2);
		up(q.D, S.qt[o].D.data(), S.qt[o].D.size() * 8);
		up(q.prow, S.pt[o].row_q.data(), S.pt[o].row_q.size() * 8);
		up(q.pcol, S.pt[o].col_mont.data(), S.pt[o].col_mont.size() * 8);
		up(q.M, S.et[o].M.data(), S.et[o].M.size() * 8);
	}
	if (!S.pack.empty())
		ok &= hipMemcpyAsync(base + t0, S.pack.data(), S.pack.size(), hipMemcpyHostToDevice, st) == hipSuccess;
	return ok;
}

// Output rows per thread of vband_kernel: the most (8, 4, 2, 1) whose bands
// reach at most BAND_MAX_K input positions.
int vband_rows(const iwamd_plan &p, const iwx::DenseWeights &d)
{
	for (int R : {8, 4, 2, 1}) {
		bool ok = true;
		for (int j0 = 0; j0 < p.inter_height && ok; j0 += R) {
			int lo, n;
			dense_span(d, j0, std::min(R, p.inter_height - j0), lo, n);
			ok = n <= BAND_MAX_K;
		}
		if (ok) return R;
	}
	return 0;
}

// Kernel arguments of a staged job placed at base.
// Pixels each segment's walk takes before its own first pixel.  Walks with
// equal levels close their gap by the in-row weights at every pixel, so the
// lead that makes a segment's start state the true one is set by the
// weights (slower with an x+2 carry) and by how often the walks' levels
// differ (more often with few, widely spaced levels); a start state that is
// still off costs a serial walk to where it meets the speculative one.
int ed_warm(const iwamd_ochan &o, int P)
{
	const int sub = (o.dithersubtype >= 0 && o.dithersubtype <= 7) ? o.dithersubtype : 0;
	const bool m1 = sub != 0 && sub != 6;
	if (P < 16) return 128;
	return m1 ? 96 : 64;
}

void fill_dev_plan(const Staged &S, const iwamd_plan &p, char *base, DevPlan &D)
{
	memset(&D, 0, sizeof(D));
	D.in = (const uint8_t *)(base + S.o_in);
	D.in_bpr = S.in_f64.empty() ? p.in_bpr : S.in_f64_bpr;
	D.in_f64 = !S.in_f64.empty();
	D.in_w = p.in_width;
	D.in_h = p.in_height;
	D.bit_depth = p.in_bit_depth;
	D.in_float = !S.in_int;
	D.nch_phys = p.in_nch_physical;
	D.orient = p.orient_transform;
	D.sx = p.input_start_x;
	D.sy = p.input_start_y;
	D.iw = p.input_w;
	D.ih = p.input_h;
	D.alpha_in_index = p.in_has_alpha ? p.in_alpha_index : 1000;
	D.alpha_lut = S.alpha_lut >= 0 ? (const double *)(base + S.o_lut[S.alpha_lut]) : nullptr;
	D.alpha_max = (p.in_has_alpha && p.in_alpha_index < 4) ? p.in_ch_maxcolorcode[p.in_alpha_index] : 0.0;
	D.gsf = p.grayscale_formula;
	for (int k = 0; k < 3; k++) D.gsw[k] = p.grayscale_weight[k];
	D.inter_nch = p.inter_nch;
	D.inter_alpha = p.inter_alpha_index;
	{
		int n = 0;
		if (p.inter_alpha_index >= 0) D.order[n++] = p.inter_alpha_index;
		for (int c = 0; c < p.inter_nch; c++)
			if (c != p.inter_alpha_index) D.order[n++] = c;
	}
	D.W = p.inter_width;
	D.Hp = p.inter_height;
	D.v_shared = (int)S.dv.size() == p.inter_nch && p.inter_nch > 0;
	for (int c = 0; c < (int)S.dv.size() && D.v_shared; c++)
		if (!same_dense(S.dv[c], S.dv[0]) || S.wv[c].mode != iwx::WeightList::WEIGHTS) D.v_shared = 0;
	D.h_shared = (int)S.dh.size() == p.inter_nch && p.inter_nch > 0;
	for (int c = 0; c < (int)S.dh.size() && D.h_shared; c++)
		if (!same_dense(S.dh[c], S.dh[0]) || S.wh[c].mode != iwx::WeightList::WEIGHTS) D.h_shared = 0;
	for (int c = 0; c < p.inter_nch; c++) {
		const iwamd_ichan &s = p.ich[c];
		DevIChan &d = D.ic[c];
		d.input_channel = s.input_channel;
		d.cvt_gray = s.cvt_to_grayscale;
		d.output_channel = s.output_channel;
		d.need_unassoc = s.need_unassoc_alpha;
		d.early_bkgd = s.early_bkgd;
		d.bkgd = s.bkgd_color_lin;
		for (int k = 0; k < 3; k++)
			d.lut[k] = (S.in_int && S.ic_lut[c][k] >= 0) ? (const double *)(base + S.o_lut[S.ic_lut[c][k]]) : nullptr;
		// The table of a linear channel holds v / max (build_input_lut), max as
		// stage_job chose it (the channel's own on the slow path).
		d.lut_lin_max = 0.0;
		if (d.lut[0] && s.in_cs.cstype == IW_CSTYPE_LINEAR && s.input_channel < p.in_nch_physical)
			d.lut_lin_max = p.in_ch_slow[s.input_channel] ? p.in_ch_maxcolorcode[s.input_channel] : p.in_maxcolorcode;
		auto fill = [&](const iwx::WeightList &wl, const WOff &o, DevWeights &dw, int dim) {
			dw.mode = wl.mode;
			dw.start = (const int *)(base + o.start);
			dw.src = (const int *)(base + o.src);
			dw.w = (const double *)(base + o.w);
			dw.copy_src = (const int *)(base + o.copy);
			dw.edge = wl.edge_value;
			dw.edge_transparent = p.rs[dim].edge_policy == IW_EDGE_POLICY_TRANSPARENT;
		};
		fill(S.wv[c], S.ov[c], d.v, IW_DIMENSION_V);
		fill(S.wh[c], S.oh[c], d.h, IW_DIMENSION_H);
		d.v.dfirst = (const int *)(base + S.ov[c].dfirst);
		d.v.dcount = (const int *)(base + S.ov[c].dcount);
		d.v.doff = (const int *)(base + S.ov[c].doff);
		d.v.dw = (const double *)(base + S.ov[c].dw);
		d.v.dmax = S.dv[c].max_count;
		d.h.dfirst = (const int *)(base + S.oh[c].dfirst);
		d.h.dcount = (const int *)(base + S.oh[c].dcount);
		d.h.doff = (const int *)(base + S.oh[c].doff);
		d.h.dw = (const double *)(base + S.oh[c].dw);
		d.h.dmax = S.dh[c].max_count;
	}
	D.inter = (float *)(base + S.o_inter);
	D.h4i = S.h4i.empty() ? nullptr : (const int4 *)(base + S.o_h4i);
	D.h4w = S.h4w.empty() ? nullptr : (const float *)(base + S.o_h4w);
	D.intclamp = p.intclamp;
	D.out_w = p.out_width;
	D.out_h = p.out_height;
	D.out_depth = p.out_bit_depth;
	D.out_float = p.out_sampletype == IW_SAMPLETYPE_FLOATINGPOINT;
	D.out_nch = p.out_nch;
	D.out_bpr = p.out_bpr;
	D.out = (uint8_t *)(base + S.o_out);
	for (int o = 0; o < p.out_nch; o++) {
		const iwamd_ochan &s = p.och[o];
		DevOChan &d = D.oc[o];
		d.family = s.ditherfamily;
		d.subtype = s.dithersubtype;
		d.is_alpha = s.channeltype == IW_CHANNELTYPE_ALPHA;
		d.ed_index = S.ed_of[o];
		d.bk1 = s.bkgd1_color_lin;
		d.bk2 = s.bkgd2_color_lin;
		if (!D.out_float) {
			const QOff &q = S.oq[o];
			d.qmode = s.use_nc_table ? Q_NC : Q_THRESH;
			d.P = S.qt[o].P;
			d.T = (const double *)(base + q.T);
			d.U = (const double *)(base + q.U);
			d.lin = (const double *)(base + q.lin);
			d.code = (const double *)(base + q.code);
			d.nc = (const double *)(base + q.nc);
			d.ix = S.qt[o].ix.empty() ? nullptr : (const unsigned short *)(base + q.ix);
			d.D = S.qt[o].D.empty() ? nullptr : (const double *)(base + q.D);
			d.code_is_level = S.qt[o].code_is_level ? 1 : 0;
			d.ix_le2 = !S.qt[o].ix.empty() && S.qt[o].ix_max <= 2;
			d.ix_log2 = S.qt[o].ix_log2;
			d.prng_row = (const uint64_t *)(base + q.prow);
			d.prng_col = (const uint64_t *)(base + q.pcol);
			d.M = (const double *)(base + q.M);
			d.ed_warm = S.ed_warm > 0 ? S.ed_warm : ed_warm(s, S.qt[o].P);
		}
	}
	D.ed_samples = (double *)(base + S.o_eds);
	D.ed_err = (double *)(base + S.o_ede);
	D.n_ed = S.n_ed;
	D.ed_seg = S.ed_seg;
	D.apply_bkgd = p.apply_bkgd;
	D.late = p.bkgd_strategy_late;
	D.checker = p.bkgd_checkerboard;
	D.check_size = p.bkgd_check_size;
	D.ox = p.bkgd_check_origin[0];
	D.oy = p.bkgd_check_origin[1];
	D.bk_has_trans = p.bkgd_has_transparency;
	D.bk1a = p.bkgd1alpha;
	D.bk2a = p.bkgd2alpha;
	D.vband_r = 0;
	FastCin F;
	if (!S.band_rows && S.fast.empty() && D.v_shared && fast_cin_depth(D, F)) D.vband_r = vband_rows(p, S.dv[0]);
}

// ---- work sets ---------------------------------------------------------------
// Device work lists for a set of staged jobs: the plan array (the descriptor
// table the kernels index), the bands of each band height, the
// error-diffusion planes of each kernel variant, and the jobs that take the
// two-pass kernels.

constexpr int kBandR[3] = {1, 2, 4};
// Bands are also grouped by the LDS they need (<= 16, 32, 64, 128 KiB, more),
// one persistent launch per group, so a batch's widest rows do not lower the
// occupancy of its narrow ones.
constexpr int kLdsClasses = 5;
// Persistent work counters: one per band list (band height r, LDS class q)
// at r * kLdsClasses + q, then XCD_COUNT per chosen-form variant.
constexpr int kCounters = 3 * kLdsClasses + XCD_COUNT * FAST_NVAR;

int lds_class(size_t bytes)
{
	int q = 0;
	while (q < kLdsClasses - 1 && bytes > ((size_t)16384 << q)) q++;
	return q;
}

typedef void (*EdKernel)(const DevPlan *, const int2 *);
constexpr int kEdVariants = 12;
const EdKernel kEdKernels[kEdVariants] = {
	errdiff_kernel<0, false, false>, errdiff_kernel<0, false, true>,
	errdiff_kernel<0, true, false>, errdiff_kernel<0, true, true>,
	errdiff_kernel<1, false, false>, errdiff_kernel<1, false, true>,
	errdiff_kernel<1, true, false>, errdiff_kernel<1, true, true>,
	errdiff_kernel<2, false, false>, errdiff_kernel<2, false, true>,
	errdiff_kernel<2, true, false>, errdiff_kernel<2, true, true>};

// Kernel variant of one error-diffusion plane: decision records in registers
// for at most ED_SMALL_P + 1 levels, in LDS for up to 257, the x+2 carry only
// where its weight is nonzero (not Floyd-Steinberg or Sierra 4-2-1), working
// rows in LDS up to ED_ROWS_LDS wide.
int ed_variant(const Staged &S, const iwamd_plan &p, int o)
{
	const int subtype = p.och[o].dithersubtype;
	const int sub = (subtype >= 0 && subtype <= 7) ? subtype : 0;
	const bool m1 = sub != 0 && sub != 6;
	const bool rows = p.out_width <= ED_ROWS_LDS && ed_cap(p.out_width, S.ed_seg) <= ED_CAP_LDS;
	const int P = S.qt[o].P;
	const int tabm = P <= ED_SMALL_P ? 2 : (P <= ED_LDS_TABLE ? 1 : 0);
	return 4 * tabm + (m1 ? 2 : 0) + (rows ? 1 : 0);
}

struct WorkSet {
	std::vector<DevPlan> plans;       // host copies; the two-pass kernels take them by value
	std::vector<int2> bands[3][kLdsClasses];  // (plan, first row), by band height kBandR and LDS class
	size_t band_lds[3][kLdsClasses] = {};      // largest LDS of a band in each list
	std::vector<int2> planes[kEdVariants];  // (plan, output channel), by ed_variant
	std::vector<int> twopass;         // plans resampled by vpass + hpass
	// Chosen-precision tiles, and work items over them: runs of one band's
	// tiles (first tile, count), by kernel variant.
	std::vector<FastItem> fast_tiles;
	std::vector<int4> fast_runs;      // (first, count, variant, -) per band, before splitting
	std::vector<int2> fast[FAST_NVAR];
	int fast_chunks[FAST_NVAR][XCD_COUNT + 1] = {};  // each chosen-form list's chunks, one per XCD
	size_t o_plans = 0, o_bands[3][kLdsClasses] = {}, o_planes[kEdVariants] = {0}, o_counters = 0, o_fast[FAST_NVAR] = {},
		o_chunks = 0, o_done = 0;
	size_t bytes = 0;
	const FastItem *d_fast_tiles = nullptr;
	const int2 *d_fast[FAST_NVAR] = {};
	size_t o_fast_tiles = 0;
	int *d_done = nullptr;  // one error-diffusion completion counter per plan
	const DevPlan *d_plans = nullptr;
	const int2 *d_bands[3][kLdsClasses] = {};
	const int2 *d_planes[kEdVariants] = {nullptr};
	const int *d_chunks = nullptr;
	int *d_counters = nullptr;
};

// The chosen-form kernel variant of a plan's tiles (FAST_*), decided by the
// kernels' own tests (fast_cin_depth, emit_simple_ctx) on the plan's kernel
// arguments as placed at a stand-in base: which tables exist and the byte
// layout of the input do not depend on where the block lands.  (Whether
// 8-bit pixels load as 32-bit words, which does, FAST_LEAN8 decides per
// item.)
int fast_variant(const Staged &S, const iwamd_plan &p)
{
	DevPlan D;
	fill_dev_plan(S, p, (char *)(uintptr_t)0x100000000ull, D);
	FastCin F;
	EmitSimple ES;
	const int depth = fast_cin_depth(D, F);
	if (!depth || !F.lean || !emit_simple_ctx(D, ES)) return FAST_GENERIC;
	if (depth == 16 || depth == 64) return FAST_LEAN16;
	return fast_rgba8_ok(D, ES, F) ? FAST_RGBA8 : FAST_LEAN8;
}

// Adds a job's work; its kernel arguments are set later (W.plans[index]),
// once its device data has a place.  Returns the job's index.
int ws_add(WorkSet &W, const Staged &S, const iwamd_plan &p)
{
	const int jx = (int)W.plans.size();
	W.plans.push_back(DevPlan());
	if (S.skip_pixels) return jx;
	if (!S.fast.empty()) {
		const int var = fast_variant(S, p);
		for (size_t k = 0; k < S.fast.size();) {
			size_t e = k;
			while (e < S.fast.size() && S.fast[e].j0 == S.fast[k].j0) e++;
			const int first = (int)W.fast_tiles.size();
			for (size_t u = k; u < e; u++) {
				FastItem it = S.fast[u];
				it.job = jx;
				W.fast_tiles.push_back(it);
			}
			W.fast_runs.push_back(make_int4(first, (int)(e - k), var, 0));
			k = e;
		}
	} else if (S.band_rows) {
		const int ri = S.band_rows == 4 ? 2 : (S.band_rows == 2 ? 1 : 0);
		const size_t lds = band_lds_bytes(S.band_rows, p.inter_nch, p.inter_width);
		const int q = lds_class(lds);
		for (int j0 = 0; j0 < p.inter_height; j0 += S.band_rows) W.bands[ri][q].push_back(make_int2(jx, j0));
		W.band_lds[ri][q] = std::max(W.band_lds[ri][q], lds);
	} else {
		W.twopass.push_back(jx);
	}
	for (int e = 0; e < S.n_ed; e++) {
		const int o = S.ed_list[e];
		W.planes[ed_variant(S, p, o)].push_back(make_int2(jx, o));
	}
	return jx;
}

int device_cu_count();

void ws_layout(WorkSet &W)
{
	// Work items of the chosen form: a band's tiles share its weights in LDS,
	// in runs of up to FAST_RUN tiles (a wave takes every fourth), fewer when
	// the work set holds too few tiles to give every workgroup slot several
	// items.
	const int per = std::max(1, std::min(FAST_RUN, (int)(W.fast_tiles.size() / (8 * (size_t)device_cu_count()))));
	for (auto &l : W.fast) l.clear();
	for (const int4 &r : W.fast_runs)
		for (int a = 0; a < r.y; a += per) W.fast[r.z].push_back(make_int2(r.x + a, std::min(per, r.y - a)));
	// Each chosen-form list cut into XCD_COUNT chunks of about equal cost, in
	// list order (a job's bands stay neighbours): a tile costs about its
	// band's input rows plus a fixed part (its horizontal products and
	// per-pixel work, counted as 32 rows).
	for (int v = 0; v < FAST_NVAR; v++) {
		const std::vector<int2> &l = W.fast[v];
		auto wt = [&](const int2 &it) { return (size_t)it.y * (size_t)(W.fast_tiles[it.x].kv + 32); };
		size_t total = 0;
		for (const int2 &it : l) total += wt(it);
		int *ch = W.fast_chunks[v];
		size_t acc = 0;
		int x = 1;
		ch[0] = 0;
		for (size_t k = 0; k < l.size() && x < XCD_COUNT; k++) {
			acc += wt(l[k]);
			while (x < XCD_COUNT && acc * XCD_COUNT >= total * (size_t)x) ch[x++] = (int)k + 1;
		}
		while (x <= XCD_COUNT) ch[x++] = (int)l.size();
	}
	Layout L;
	W.o_plans = L.add<DevPlan>(W.plans.size());
	for (int r = 0; r < 3; r++)
		for (int q = 0; q < kLdsClasses; q++) W.o_bands[r][q] = L.add<int2>(W.bands[r][q].size());
	for (int v = 0; v < kEdVariants; v++) W.o_planes[v] = L.add<int2>(W.planes[v].size());
	W.o_counters = L.add<int>(kCounters);
	W.o_fast_tiles = L.add<FastItem>(W.fast_tiles.size());
	for (int v = 0; v < FAST_NVAR; v++) W.o_fast[v] = L.add<int2>(W.fast[v].size());
	W.o_chunks = L.add<int>(FAST_NVAR * (XCD_COUNT + 1));
	W.o_done = L.add<int>(W.plans.size());
	W.bytes = (L.off + 255) & ~(size_t)255;
}

// Queue the copies of the work lists to base (W must live until the stream
// is synchronised).
bool ws_upload(WorkSet &W, char *base, hipStream_t st)
{
	bool ok = true;
	auto up = [&](size_t off, const void *src, size_t bytes) {
		if (bytes) ok &= hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
	};
	W.d_plans = (const DevPlan *)(base + W.o_plans);
	up(W.o_plans, W.plans.data(), W.plans.size() * sizeof(DevPlan));
	for (int r = 0; r < 3; r++) {
		for (int q = 0; q < kLdsClasses; q++) {
			W.d_bands[r][q] = (const int2 *)(base + W.o_bands[r][q]);
			up(W.o_bands[r][q], W.bands[r][q].data(), W.bands[r][q].size() * sizeof(int2));
		}
	}
	for (int v = 0; v < kEdVariants; v++) {
		W.d_planes[v] = (const int2 *)(base + W.o_planes[v]);
		up(W.o_planes[v], W.planes[v].data(), W.planes[v].size() * sizeof(int2));
	}
	W.d_counters = (int *)(base + W.o_counters);
	W.d_fast_tiles = (const FastItem *)(base + W.o_fast_tiles);
	up(W.o_fast_tiles, W.fast_tiles.data(), W.fast_tiles.size() * sizeof(FastItem));
	for (int v = 0; v < FAST_NVAR; v++) {
		W.d_fast[v] = (const int2 *)(base + W.o_fast[v]);
		up(W.o_fast[v], W.fast[v].data(), W.fast[v].size() * sizeof(int2));
	}
	W.d_chunks = (const int *)(base + W.o_chunks);
	up(W.o_chunks, W.fast_chunks, sizeof(W.fast_chunks));
	W.d_done = (int *)(base + W.o_done);
	return ok;
}

int device_cu_count()
{
	static int n = 0;
	if (!n) {
		int dev = 0, v = 0;
		if (hipGetDevice(&dev) == hipSuccess &&
			hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
			n = v;
		else
			n = 256;
	}
	return n;
}

// Queue the resampling kernels of a work set.  persistent: one launch per
// band height sized to the device, workgroups taking bands from a counter;
// otherwise one workgroup per band.
void ws_enqueue_resample(const WorkSet &W, hipStream_t st, bool persistent)
{
	if (persistent) {
		hipLaunchKernelGGL(reset_counters_kernel, dim3((kCounters + 63) / 64), dim3(64), 0, st, W.d_counters,
			kCounters);
		const int np = (int)W.plans.size();
		hipLaunchKernelGGL(reset_counters_kernel, dim3((np + 255) / 256), dim3(256), 0, st, W.d_done, np);
	}
	for (int v = 0; v < FAST_NVAR; v++) {
		const int n = (int)W.fast[v].size();
		if (!n) continue;
		const int grid = persistent ? std::min(n, device_cu_count() * (v == FAST_GENERIC ? 1 : FAST_WG_PER_CU)) : n;
		int *ctr = persistent ? W.d_counters + 3 * kLdsClasses + XCD_COUNT * v : nullptr;
		const int *ch = W.d_chunks + v * (XCD_COUNT + 1);
		switch (v) {
		case FAST_RGBA8:
			hipLaunchKernelGGL(fast_kernel<FAST_RGBA8>, dim3(grid), dim3(256), 0, st, W.d_plans, W.d_fast_tiles,
				W.d_fast[v], n, ctr, ch);
			break;
		case FAST_LEAN8:
			hipLaunchKernelGGL(fast_kernel<FAST_LEAN8>, dim3(grid), dim3(256), 0, st, W.d_plans, W.d_fast_tiles,
				W.d_fast[v], n, ctr, ch);
			break;
		case FAST_LEAN16:
			hipLaunchKernelGGL(fast_kernel<FAST_LEAN16>, dim3(grid), dim3(256), 0, st, W.d_plans, W.d_fast_tiles,
				W.d_fast[v], n, ctr, ch);
			break;
		default:
			hipLaunchKernelGGL(fast_kernel<FAST_GENERIC>, dim3(grid), dim3(256), 0, st, W.d_plans, W.d_fast_tiles,
				W.d_fast[v], n, ctr, ch);
			break;
		}
	}
	for (int r = 0; r < 3; r++)
		for (int q = 0; q < kLdsClasses; q++) {
			const int n = (int)W.bands[r][q].size();
			if (!n) continue;
			const size_t lds = W.band_lds[r][q];
			const int2 *items = W.d_bands[r][q];
			int grid = n;
			if (persistent) {
				const int per_cu = std::max(1, std::min(8, (int)(rt().band_lds_max / std::max<size_t>(lds, 1))));
				grid = std::min(n, device_cu_count() * per_cu);
			}
			int *ctr = persistent ? W.d_counters + r * kLdsClasses + q : nullptr;
			switch (kBandR[r]) {
			case 1: hipLaunchKernelGGL(band_kernel<1>, dim3(grid), dim3(256), lds, st, W.d_plans, items, n, ctr); break;
			case 2: hipLaunchKernelGGL(band_kernel<2>, dim3(grid), dim3(256), lds, st, W.d_plans, items, n, ctr); break;
			default: hipLaunchKernelGGL(band_kernel<4>, dim3(grid), dim3(256), lds, st, W.d_plans, items, n, ctr); break;
			}
		}
	for (int jx : W.twopass) {
		const DevPlan &D = W.plans[jx];
		enqueue_vpass(D, st);
		hipLaunchKernelGGL(hpass_kernel, dim3((D.out_w + 255) / 256, D.out_h), dim3(256), 0, st, D);
	}
}

void ws_enqueue_errdiff(const WorkSet &W, hipStream_t st)
{
	for (int v = 0; v < kEdVariants; v++) {
		const int n = (int)W.planes[v].size();
		if (n) hipLaunchKernelGGL(kEdKernels[v], dim3(n), dim3(ED_THREADS), 0, st, W.d_plans, W.d_planes[v]);
	}
}

// Useful floating-point operations of a plan's resampling: a multiply and an
// add per listed tap, vertically for every intermediate sample and
// horizontally for every output sample, per channel.
double resample_flops(const iwamd_plan &p, const Staged &S)
{
	if (S.skip_pixels || S.dv.empty() || S.dh.empty()) return 0.0;
	double f = 0.0;
	for (int c = 0; c < p.inter_nch; c++) {
		double tv = 0.0, th = 0.0;
		for (int j = 0; j < p.inter_height; j++) tv += S.dv[c].count[j];
		for (int i = 0; i < p.out_width; i++) th += S.dh[c].count[i];
		f += 2.0 * (tv * p.inter_width + th * p.out_height);
	}
	return f;
}

// ---- autotuner --------------------------------------------------------------
//
// The band height of the band kernel (4, 2 or 1 output rows per workgroup)
// and the tile width of the chosen-precision form (32 or 16 output columns)
// change how the work is cut, not the bytes: a band row's accumulator gets
// its list's terms in list order plus signed zeros for the band's other
// positions, which leave it unchanged, and a tile's 16-output groups and the
// 16-column chunks they sum over do not depend on the tile width.  So the
// shape for a class of jobs -- form, channels, vertical and horizontal taps
// per output (filter radius x scale factor) and output width, each as
// ceil(log2) -- is chosen by timing: the first job of a class that may be
// tuned runs its resampling once per candidate shape on its own device copy,
// its work repeated to fill the device (a warm run, then the best of three;
// time per copy: a batch's throughput, not one job's latency), the fastest
// shape is kept for the class, and later jobs of the class take it.  Batches tune; single jobs
// and the request loop only look the class up (amd:tune=1 lets a single job
// tune; IWAMD_TUNE=0 in the environment turns tuning off).

struct TuneKey {
	int form, nch, tv, th, wb;
	bool operator<(const TuneKey &o) const
	{
		return std::tie(form, nch, tv, th, wb) < std::tie(o.form, o.nch, o.tv, o.th, o.wb);
	}
};

struct TuneEntry {
	int shape = 0, n = 0;
	int cand[3] = {0, 0, 0};
	double ms[3] = {0, 0, 0};
};

struct Tuner {
	std::mutex mu;
	std::map<TuneKey, TuneEntry> table;
};

Tuner &tuner()
{
	static Tuner t;
	return t;
}

int log2_bucket(double v)
{
	int b = 0;
	while (b < 30 && (double)(1 << b) < v) b++;
	return b;
}

// 0: band kernel, 1: chosen-precision tiles, -1: nothing to tune.
int tune_form(const Staged &S)
{
	if (S.skip_pixels) return -1;
	if (!S.fast.empty()) return 1;
	return S.band_rows ? 0 : -1;
}

TuneKey tune_key(const iwamd_plan &p, const Staged &S, int form)
{
	auto mean_taps = [](const iwx::DenseWeights &d, int n) {
		double t = 0.0;
		for (int k = 0; k < n; k++) t += d.count[k];
		return n ? t / n : 0.0;
	};
	TuneKey k;
	k.form = form;
	k.nch = p.inter_nch;
	k.tv = log2_bucket(mean_taps(S.dv[0], p.inter_height));
	k.th = log2_bucket(mean_taps(S.dh[0], p.out_width));
	k.wb = log2_bucket(p.out_width / 64.0);
	return k;
}

int tune_cands(const iwamd_plan &p, const Staged &S, int form, int *cand)
{
	int n = 0;
	if (form == 0) {
		for (int R : {4, 2, 1})
			if (band_fits(p, S.dv, R)) cand[n++] = R;
	} else {
		cand[n++] = 32;
		cand[n++] = 16;
	}
	return n;
}

void tune_set(const iwamd_plan &p, Staged &S, int form, int shape)
{
	if (form == 0) S.band_rows = shape;
	else plan_fast(p, S, shape);
}

// One job's resampling cost with each candidate shape, on its own device
// copy, as a batch would see it: the job's work items are repeated until the
// list holds about four workgroups per CU (the copies write the same bytes to
// the same output), and the time is divided by the number of copies.
bool tune_measure(const iwamd_plan &p, Staged &S, int form, const int *cand, int n, double *ms, hipStream_t st)
{
	const size_t out_al = (S.out_bytes + 255) & ~(size_t)255;
	void *blk = dev_alloc(S.bytes + out_al);
	if (!blk) return false;
	char *base = (char *)blk;
	bool ok = upload_job(S, p, base, st);
	hipEvent_t e0 = nullptr, e1 = nullptr;
	ok = ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
	for (int k = 0; k < n && ok; k++) {
		tune_set(p, S, form, cand[k]);
		const int items = form == 0 ? (p.inter_height + S.band_rows - 1) / S.band_rows : (int)S.fast.size();
		const int copies = std::max(1, std::min(256, 4 * device_cu_count() / std::max(1, items)));
		WorkSet W;
		DevPlan D;
		fill_dev_plan(S, p, base, D);
		D.out = (uint8_t *)(base + S.bytes);
		for (int c = 0; c < copies; c++) {
			ws_add(W, S, p);
			W.plans.back() = D;
		}
		ws_layout(W);
		void *wb = dev_alloc(W.bytes);
		if (!wb) {
			ok = false;
			break;
		}
		ok = ws_upload(W, (char *)wb, st);
		double best = 1e30;
		for (int r = 0; r < 4 && ok; r++) {
			ok = hipEventRecord(e0, st) == hipSuccess;
			ws_enqueue_resample(W, st, true);
			ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
			float t = 0.0f;
			if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess && r > 0) best = std::min(best, (double)t);
		}
		ms[k] = best / copies;
		ok = ok && hipStreamSynchronize(st) == hipSuccess;
		dev_release(wb, W.bytes);
	}
	if (e0) hipEventDestroy(e0);
	if (e1) hipEventDestroy(e1);
	(void)hipStreamSynchronize(st);
	dev_release(blk, S.bytes + out_al);
	return ok;
}

bool tune_enabled()
{
	const char *v = getenv("IWAMD_TUNE");
	return !(v && v[0] == '0');
}

// Gives the plan its class's tuned shape; a class without one is measured on
// this job first when may_tune.  A shape that does not fit this job (or a
// failed measurement) keeps the default choice.
void tune_apply(const iwamd_plan &p, Staged &S, hipStream_t st, bool may_tune)
{
	const int form = tune_form(S);
	if (form < 0 || S.shape_forced) return;
	int cand[3];
	const int n = tune_cands(p, S, form, cand);
	if (n < 2) return;
	const int orig = form == 0 ? S.band_rows : S.fast_ti;
	const TuneKey key = tune_key(p, S, form);
	Tuner &T = tuner();
	int shape = 0;
	{
		std::lock_guard<std::mutex> lock(T.mu);
		auto it = T.table.find(key);
		if (it != T.table.end()) shape = it->second.shape;
	}
	if (!shape && may_tune && tune_enabled()) {
		TuneEntry e;
		if (tune_measure(p, S, form, cand, n, e.ms, st)) {
			int best = 0;
			for (int k = 1; k < n; k++)
				if (e.ms[k] < e.ms[best]) best = k;
			e.shape = cand[best];
			e.n = n;
			for (int k = 0; k < n; k++) e.cand[k] = cand[k];
			std::lock_guard<std::mutex> lock(T.mu);
			shape = T.table.emplace(key, e).first->second.shape;
		}
	}
	bool fits = false;
	for (int k = 0; k < n; k++) fits |= shape && cand[k] == shape;
	tune_set(p, S, form, fits ? shape : orig);
}

bool job_upload(iwamd_job *J)
{
	iw_context *ctx = J->ctx;
	const iwamd_plan &p = J->plan;
	Staged S;
	if (!stage_job(ctx, p, S, iw_get_option(ctx, "amd:resample"))) return false;
	{
		const char *t = iw_get_option(ctx, "amd:tune");
		tune_apply(p, S, J->stream, t && atoi(t) == 1);
	}
	J->skip_pixels = S.skip_pixels;
	J->n_ed = S.n_ed;
	for (int e = 0; e < 4; e++) J->ed_list[e] = S.ed_list[e];
	J->band_rows = S.band_rows;
	J->out_off = S.o_out;
	J->out_bytes = S.out_bytes;
	// Work lists follow the job's data in the same block; their sizes do not
	// depend on where the block lands.
	WorkSet *W = new WorkSet();
	J->ws = W;
	ws_add(*W, S, p);
	ws_layout(*W);
	J->block_bytes = S.bytes + W->bytes;
	J->block = dev_alloc(J->block_bytes);
	if (!J->block) {
		iw_set_error(ctx, "Out of device memory");
		return false;
	}
	char *base = (char *)J->block;
	fill_dev_plan(S, p, base, J->D);
	W->plans[0] = J->D;
	bool ok = upload_job(S, p, base, J->stream);
	ok &= ws_upload(*W, base + S.bytes, J->stream);
	// Host tables are freed when S goes out of scope: wait for the copies.
	ok &= hipStreamSynchronize(J->stream) == hipSuccess;
	if (!ok) {
		iw_set_error(ctx, "Host-to-device copy failed");
		return false;
	}
	J->out_pixels = (double)p.out_width * p.out_height;
	J->alg_bytes = (double)S.in_bytes + (double)S.out_bytes;
	J->flops = resample_flops(p, S);
	return true;
}

}  // namespace

extern "C" struct iwamd_job *iwamd_job_create(struct iw_context *ctx)
{
	iwamd_job *J = new iwamd_job();
	J->ctx = ctx;
	// Host decisions first, so their errors read as the reference's do.
	if (!iwamd_prepare(ctx, &J->plan) || !runtime_init(ctx)) {
		delete J;
		return nullptr;
	}
	if (hipStreamCreateWithFlags(&J->stream, hipStreamNonBlocking) != hipSuccess) {
		iw_set_error(ctx, "hipStreamCreate failed");
		delete J;
		return nullptr;
	}
	for (auto &e : J->ev) hipEventCreate(&e);
	if (!job_upload(J)) {
		iwamd_job_destroy(J);
		return nullptr;
	}
	return J;
}

namespace {

// One job's kernels on st (one workgroup per band, not persistent); with
// events, ev[0..3] bracket the resampling, the horizontal pass (empty when
// fused) and error diffusion.
void enqueue_job(const WorkSet &W, const DevPlan &D, hipStream_t st, hipEvent_t *ev)
{
	if (ev) hipEventRecord(ev[0], st);
	if (W.twopass.empty()) {
		// One fused kernel; the horizontal-pass slot of the timings stays 0.
		ws_enqueue_resample(W, st, false);
		if (ev) hipEventRecord(ev[1], st);
	} else {
		enqueue_vpass(D, st);
		if (ev) hipEventRecord(ev[1], st);
		hipLaunchKernelGGL(hpass_kernel, dim3((D.out_w + 255) / 256, D.out_h), dim3(256), 0, st, D);
	}
	if (ev) hipEventRecord(ev[2], st);
	ws_enqueue_errdiff(W, st);
	if (ev) hipEventRecord(ev[3], st);
}

}  // namespace

extern "C" int iwamd_job_run(struct iwamd_job *J)
{
	if (J->skip_pixels) return 1;
	enqueue_job(*(const WorkSet *)J->ws, J->D, J->stream, J->ev);
	if (hipGetLastError() != hipSuccess) {
		iw_set_error(J->ctx, "Kernel launch failed");
		return 0;
	}
	J->ran = true;
	return 1;
}

extern "C" int iwamd_job_sync(struct iwamd_job *J)
{
	hipError_t e = hipStreamSynchronize(J->stream);
	if (e != hipSuccess) {
		iw_set_errorf(J->ctx, "GPU pipeline failed: %s", hipGetErrorString(e));
		return 0;
	}
	return 1;
}

extern "C" int iwamd_job_kernel_ms(struct iwamd_job *J, double *ms, int n)
{
	if (!J->ran) return 0;
	// A fused resample launches nothing between the first two events: its
	// horizontal-pass slot is 0 (the events alone read a few microseconds).
	const bool fused = ((const WorkSet *)J->ws)->twopass.empty();
	int k = 0;
	for (; k < 3 && k < n; k++) {
		float t = 0.0f;
		if (!(k == 1 && fused)) hipEventElapsedTime(&t, J->ev[k], J->ev[k + 1]);
		ms[k] = t;
	}
	return k;
}

extern "C" void iwamd_job_sizes(struct iwamd_job *J, double *out_pixels, double *alg_bytes)
{
	*out_pixels = J->out_pixels;
	*alg_bytes = J->alg_bytes;
}

extern "C" double iwamd_job_flops(struct iwamd_job *J)
{
	return J->flops;
}

extern "C" int iwamd_job_finish(struct iwamd_job *J)
{
	std::vector<uint8_t> host(J->out_bytes + 1, 0);
	// The optimiser's scans run on the output before it is copied back.
	const unsigned int want = J->skip_pixels ? 0 : opt_scan_want(J->ctx);
	const size_t sbytes = opt_scan_bytes(1);
	void *sblk = want ? dev_alloc(sbytes) : nullptr;
	std::vector<OptJob> scans;
	std::unique_ptr<iwx_opt_scan> scan(want ? new iwx_opt_scan() : nullptr);
	bool ok = !want || sblk;
	if (ok && want) {
		iwx_opt_scan *res = (iwx_opt_scan *)((char *)sblk + opt_scan_bytes(1) - sizeof(iwx_opt_scan));
		scans.push_back(opt_job(J->ctx, (char *)J->block + J->out_off, res, want));
		ok = opt_scan_enqueue(scans, sblk, 1, J->stream) &&
			hipMemcpyAsync(scan.get(), res, sizeof(iwx_opt_scan), hipMemcpyDeviceToHost, J->stream) == hipSuccess;
	}
	if (ok && !J->skip_pixels)
		ok = hipMemcpyAsync(host.data(), (char *)J->block + J->out_off, J->out_bytes, hipMemcpyDeviceToHost,
			J->stream) == hipSuccess;
	ok = ok && iwamd_job_sync(J);
	if (sblk) dev_release(sblk, sbytes);
	if (!ok) {
		iw_set_error(J->ctx, "Device-to-host copy failed");
		return 0;
	}
	if (scan) scan->valid = want;
	double ms[3];
	int k = iwamd_job_kernel_ms(J, ms, 3);
	J->ctx->num_last_timings = k;
	for (int i = 0; i < k; i++) J->ctx->last_timings_ms[i] = ms[i];
	return iwx_finish_image(J->ctx, host.data(), scan.get());
}

extern "C" void iwamd_job_destroy(struct iwamd_job *J)
{
	if (!J) return;
	if (J->stream) hipStreamSynchronize(J->stream);
	dev_release(J->block, J->block_bytes);
	delete (WorkSet *)J->ws;
	for (auto &e : J->ev)
		if (e) hipEventDestroy(e);
	if (J->stream) hipStreamDestroy(J->stream);
	delete J;
}

// ---- a job that allocates nothing --------------------------------------------
// For callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid it).  The input image is already
// in device memory, the output goes to device memory the caller owns, and
// everything else the kernels need -- tables, the FP32 intermediate, the
// error-diffusion samples and working rows, the descriptor table and work
// lists -- lives in `scratch`, one fixed array per GPU held until the process
// ends.  Host code reaches the array only through the address that
// hipGetSymbolAddress returns on each device (the symbol's name is not a
// device address in host code).

#define IWAMD_NOALLOC_SCRATCH ((size_t)64 << 20)

static __device__ __align__(256) unsigned char scratch[IWAMD_NOALLOC_SCRATCH];

__global__ void __launch_bounds__(256) zero_bytes_kernel(uint8_t *p, size_t n)
{
	for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
		p[k] = 0;
}

// Negated samples of the non-alpha channels (IW_VAL_NEGATE_TARGET): 8- and
// 16-bit codes become maxcolorcode - code, floats 1 - s
// (src/imagew-main.c:1324-1366).
struct NegateArgs {
	unsigned int maxcc[4];
	int neg[4];
};

__global__ void __launch_bounds__(256) negate_kernel(uint8_t *img, size_t bpr, int w, int h, int nc, int depth,
	int is_float, NegateArgs a)
{
	const size_t n = (size_t)w * h * nc;
	for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
		const int ch = (int)(k % nc);
		if (!a.neg[ch]) continue;
		const size_t px = k / nc;
		uint8_t *row = img + (px / w) * bpr;
		const size_t s = (px % w) * nc + ch;
		if (is_float) {
			float f;
			memcpy(&f, row + s * 4, 4);
			f = 1.0f - f;
			memcpy(row + s * 4, &f, 4);
		} else if (depth == 8) {
			row[s] = (uint8_t)(a.maxcc[ch] - row[s]);
		} else if (depth == 16) {
			const unsigned int v = a.maxcc[ch] - (row[s * 2] * 256u + row[s * 2 + 1]);
			row[s * 2] = (uint8_t)((v & 0xff00) >> 8);
			row[s * 2 + 1] = (uint8_t)(v & 0xff);
		}
	}
}

// The colour of fully transparent pixels cleared, as the reference's output
// optimiser does for every integer image with alpha before any optional
// optimisation (src/imagew-opt.c:1112-1163, :1200): alpha is the last of nc
// channels; 16-bit samples are two bytes.
__global__ void __launch_bounds__(256) clear_transparent_kernel(uint8_t *img, size_t bpr, int w, int h, int nc,
	int bytes_per_sample)
{
	const size_t n = (size_t)w * h;
	for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
		uint8_t *px = img + (k / w) * bpr + (k % w) * nc * bytes_per_sample;
		const uint8_t *a = px + (nc - 1) * bytes_per_sample;
		if (a[0] == 0 && (bytes_per_sample == 1 || a[1] == 0))
			for (int b = 0; b < (nc - 1) * bytes_per_sample; b++) px[b] = 0;
	}
}

namespace {

// Per device: the array's address there, and the lock that makes calls on
// one device take turns (they share its array).
struct ScratchSlot {
	std::mutex mu;
	char *ptr = nullptr;
};

ScratchSlot &scratch_slot(int dev)
{
	static std::mutex table_mu;
	static std::map<int, ScratchSlot *> table;
	std::lock_guard<std::mutex> lock(table_mu);
	ScratchSlot *&s = table[dev];
	if (!s) s = new ScratchSlot();
	return *s;
}

}  // namespace

extern "C" size_t iwamd_noalloc_scratch_bytes(void)
{
	return IWAMD_NOALLOC_SCRATCH;
}

extern "C" int iwamd_process_device_noalloc(struct iw_context *ctx, const void *d_in, size_t in_bytes, void *d_out,
	size_t out_bytes, void *stream, struct iw_image *out_desc)
{
	iwamd_plan p;
	if (!iwamd_prepare(ctx, &p) || !runtime_init(ctx)) return 0;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) {
		iw_set_error(ctx, "hipGetDevice failed");
		return 0;
	}
	const int phys_rows = (p.orient_transform & 4) ? p.in_width : p.in_height;
	const size_t need_in = p.in_bpr * (size_t)phys_rows, need_out = p.out_bpr * (size_t)p.out_height;
	if (!d_in || in_bytes < need_in) {
		iw_set_errorf(ctx, "Device input: %zu bytes needed, %zu given", need_in, d_in ? in_bytes : (size_t)0);
		return 0;
	}
	if (!d_out || out_bytes < need_out) {
		iw_set_errorf(ctx, "Device output: %zu bytes needed, %zu given", need_out, d_out ? out_bytes : (size_t)0);
		return 0;
	}
	hipStream_t st = (hipStream_t)stream;
	Staged S;
	if (!stage_job(ctx, p, S, iw_get_option(ctx, "amd:resample"), true, true)) return 0;
	tune_apply(p, S, st, false);  // look the shape up; never tunes (tuning allocates)
	WorkSet W;
	ws_add(W, S, p);
	ws_layout(W);
	const size_t need = S.bytes + W.bytes;
	if (need > IWAMD_NOALLOC_SCRATCH) {
		iw_set_errorf(ctx, "This job needs %zu bytes of device scratch; iwamd_process_device_noalloc has %zu",
			need, (size_t)IWAMD_NOALLOC_SCRATCH);
		return 0;
	}
	ScratchSlot &slot = scratch_slot(dev);
	std::lock_guard<std::mutex> lock(slot.mu);
	if (!slot.ptr) {
		// Once per device, with that device current.
		void *ptr = nullptr;
		const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr) {
			iw_set_errorf(ctx, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
			return 0;
		}
		slot.ptr = (char *)ptr;
	}
	char *base = slot.ptr;
	DevPlan D;
	fill_dev_plan(S, p, base, D);
	D.in = (const uint8_t *)d_in;
	D.out = (uint8_t *)d_out;
	bool ok = true;
	if (S.skip_pixels) {
		// An unknown filter: the reference reports it and outputs zeros.
		const int grid = (int)std::min<size_t>(4096, (need_out + 255) / 256);
		hipLaunchKernelGGL(zero_bytes_kernel, dim3(std::max(grid, 1)), dim3(256), 0, st, (uint8_t *)d_out, need_out);
	} else {
		W.plans[0] = D;
		ok = upload_job(S, p, base, st) && ws_upload(W, base + S.bytes, st);
		if (ok) enqueue_job(W, D, st, nullptr);
		if (ok && ctx->req.negate_target) {
			NegateArgs a;
			memset(&a, 0, sizeof(a));
			for (int ch = 0; ch < p.out_nch && ch < 4; ch++) {
				a.neg[ch] = ctx->img2_ci[ch].channeltype != IW_CHANNELTYPE_ALPHA;
				a.maxcc[ch] = (unsigned int)ctx->img2_ci[ch].maxcolorcode_int;
			}
			const size_t ns = (size_t)p.out_width * p.out_height * p.out_nch;
			const int grid = (int)std::max<size_t>(1, std::min<size_t>(4096, (ns + 255) / 256));
			hipLaunchKernelGGL(negate_kernel, dim3(grid), dim3(256), 0, st, (uint8_t *)d_out, p.out_bpr, p.out_width,
				p.out_height, p.out_nch, p.out_bit_depth, p.out_sampletype == IW_SAMPLETYPE_FLOATINGPOINT ? 1 : 0, a);
		}
		if (ok && p.out_sampletype == IW_SAMPLETYPE_UINT && !ctx->reduced_output_maxcolor_flag &&
			(p.out_imgtype & 0x100) && (p.out_bit_depth == 8 || p.out_bit_depth == 16)) {
			const size_t npx = (size_t)p.out_width * p.out_height;
			const int grid = (int)std::max<size_t>(1, std::min<size_t>(4096, (npx + 255) / 256));
			hipLaunchKernelGGL(clear_transparent_kernel, dim3(grid), dim3(256), 0, st, (uint8_t *)d_out, p.out_bpr,
				p.out_width, p.out_height, p.out_nch, p.out_bit_depth / 8);
		}
	}
	ok = ok && hipGetLastError() == hipSuccess;
	// The host tables live in S and W: wait for the copies (and the kernels).
	const hipError_t e = hipStreamSynchronize(st);
	if (!ok || e != hipSuccess) {
		iw_set_errorf(ctx, "GPU pipeline failed: %s", hipGetErrorString(e != hipSuccess ? e : hipErrorUnknown));
		return 0;
	}
	if (out_desc) {
		memset(out_desc, 0, sizeof(*out_desc));
		out_desc->imgtype = p.out_imgtype;
		out_desc->bit_depth = p.out_bit_depth;
		out_desc->sampletype = p.out_sampletype;
		out_desc->width = p.out_width;
		out_desc->height = p.out_height;
		out_desc->bpr = p.out_bpr;
	}
	return 1;
}

// ---- HBM counter calibration ---------------------------------------------

__global__ void __launch_bounds__(256) copy_probe_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
	size_t n)
{
	for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
		dst[k] = src[k];
}

// The autotuner's table: one entry per tuned class, in key order.
extern "C" int iwamd_tune_entries(struct iwamd_tune_entry *out, int max)
{
	Tuner &T = tuner();
	std::lock_guard<std::mutex> lock(T.mu);
	int n = 0;
	for (const auto &kv : T.table) {
		if (out && n < max) {
			iwamd_tune_entry &e = out[n];
			e.form = kv.first.form;
			e.channels = kv.first.nch;
			e.taps_v_log2 = kv.first.tv;
			e.taps_h_log2 = kv.first.th;
			e.width_log2 = kv.first.wb + 6;
			e.shape = kv.second.shape;
			e.n = kv.second.n;
			for (int k = 0; k < 3; k++) {
				e.cand[k] = kv.second.cand[k];
				e.ms[k] = kv.second.ms[k];
			}
		}
		n++;
	}
	return n;
}

// ed_decide on the host, for tests: the error and level of each sample
// s[i] in [0, 1] through the same tables the walk reads -- the level records
// and bin index (tab = 1, as for up to 257 levels) or the bisection over the
// thresholds (tab = 0).  Returns 1, 0 when tab = 1 and the channel has more
// levels than the records hold, -1 for a nearest-colour plan.
extern "C" int iwamd_ed_decide_check(const struct iwamd_ochan *och, int tab, int n, const double *s, double *err,
	int *level)
{
	if (och->use_nc_table) return -1;
	const iwx::QuantTable q = iwx::build_quant_table(*och, false);
	const iwx::EdTables t = iwx::build_ed_tables(q);
	const int P = q.P;
	if (tab == 1 && P > ED_LDS_TABLE) return 0;
	if (tab == 2 && P > ED_SMALL_P) return 0;
	std::vector<unsigned short> Q(ED_NB + 2);
	for (int k = 0; k <= ED_NB; k++) Q[k] = (unsigned short)level_in(q.T.data(), 0, P, k / (double)ED_NB);
	std::vector<EdRec> rec(P + 3);
	for (int k = 0; k <= P + 2; k++) {
		if (k <= P) rec[k] = EdRec{k < P ? q.T[k] : 2.0, q.U[k], t.M[k], q.lin[k]};
		else rec[k] = EdRec{2.0, 2.0, 2.0, q.lin[P]};
	}
	EdDecide d;
	d.rec = rec.data();
	d.Q = Q.data();
	d.T = q.T.data();
	d.U = q.U.data();
	d.M = t.M.data();
	d.lin = q.lin.data();
	d.P = P;
	ed_small_fill(d.sm, P, q.T.data(), q.U.data(), t.M.data(), q.lin.data());
	for (int i = 0; i < n; i++)
		err[i] = tab == 2 ? ed_decide_small(d.sm, s[i], level[i]) :
			(tab ? ed_decide<true>(d, s[i], level[i]) : ed_decide<false>(d, s[i], level[i]));
	return 1;
}

// The no-dither quantiser as the kernels run it, on the host, for tests: the
// code of each sample from the folded boundaries QuantTable::D through the
// bin index (qbin, then the bin's entries), as count_le does on the device.
extern "C" int iwamd_quant_fold_check(const struct iwamd_ochan *och, int n, const double *s, double *code)
{
	if (och->use_nc_table) return -1;
	const iwx::QuantTable q = iwx::build_quant_table(*och, false, true);
	if (q.ix.empty()) return 0;
	const int NB = 1 << q.ix_log2;
	for (int i = 0; i < n; i++) {
		const double f = s[i] * (double)NB;
		const int b = f >= (double)(NB - 1) ? NB - 1 : (int)f;
		const int lo = q.ix[b], hi = q.ix[b + 1];
		const int k = (int)(std::upper_bound(q.D.begin() + lo, q.D.begin() + hi, s[i]) - q.D.begin());
		code[i] = q.code[k];
	}
	return 1;
}

extern "C" int iwamd_copy_probe(size_t bytes, double *ms)
{
	const size_t n = bytes / sizeof(uint4);
	if (n == 0) return 0;
	void *a = nullptr, *b = nullptr;
	if (hipMalloc(&a, n * sizeof(uint4)) != hipSuccess) return 0;
	if (hipMalloc(&b, n * sizeof(uint4)) != hipSuccess) {
		(void)hipFree(a);
		return 0;
	}
	hipEvent_t e0 = nullptr, e1 = nullptr;
	bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
	ok = ok && hipMemset(a, 0, n * sizeof(uint4)) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
	if (ok) {
		hipEventRecord(e0, 0);
		hipLaunchKernelGGL(copy_probe_kernel, dim3(4096), dim3(256), 0, 0, (const uint4 *)a, (uint4 *)b, n);
		hipEventRecord(e1, 0);
		ok = hipEventSynchronize(e1) == hipSuccess;
		float t = 0.0f;
		if (ok && ms) {
			hipEventElapsedTime(&t, e0, e1);
			*ms = t;
		}
	}
	if (e0) hipEventDestroy(e0);
	if (e1) hipEventDestroy(e1);
	(void)hipFree(a);
	(void)hipFree(b);
	return ok ? 1 : 0;
}

// ---- batches --------------------------------------------------------------
// Many contexts' pixel work in one device block: every job's input and
// tables, one output arena (one copy back), a descriptor table of kernel
// arguments, and work lists.  A run is one hipGraph: the persistent band
// kernels (one per band height, workgroups taking bands from a counter),
// the two-pass kernels of jobs too wide for a band, and the error-diffusion
// planes grouped by kernel variant.  Each job's bytes depend only on its own
// plan: band height, work order and the persistent schedule change which
// workgroup computes a band, never how.

struct iwamd_batch {
	struct Job {
		iw_context *ctx = nullptr;
		iwamd_plan plan;
		Staged S;
		size_t off = 0, out_off = 0;
		int index = -1;
		bool skip_pixels = false;
	};
	std::vector<Job *> jobs;
	WorkSet W;
	void *block = nullptr;
	size_t block_bytes = 0, out_base = 0, out_total = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev[2] = {nullptr, nullptr};
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	bool uploaded = false, ran = false, failed = false;
	double out_pixels = 0.0, alg_bytes = 0.0, flops = 0.0;
};

extern "C" struct iwamd_batch *iwamd_batch_create(void)
{
	iwamd_batch *B = new iwamd_batch();
	if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
		delete B;
		return nullptr;
	}
	for (auto &e : B->ev) hipEventCreate(&e);
	return B;
}

extern "C" int iwamd_batch_add(struct iwamd_batch *B, struct iw_context *ctx)
{
	if (B->uploaded) {
		iw_set_error(ctx, "Batch already uploaded");
		return -1;
	}
	iwamd_batch::Job *J = new iwamd_batch::Job();
	J->ctx = ctx;
	if (!iwamd_prepare(ctx, &J->plan) || !runtime_init(ctx) ||
		!stage_job(ctx, J->plan, J->S, iw_get_option(ctx, "amd:resample"), true)) {
		delete J;
		return -1;
	}
	J->index = (int)B->jobs.size();
	J->skip_pixels = J->S.skip_pixels;
	B->jobs.push_back(J);
	return J->index;
}

extern "C" int iwamd_batch_upload(struct iwamd_batch *B)
{
	if (B->uploaded) return 1;
	if (B->jobs.empty()) return 0;
	// Shapes by class (the first job of an untuned class is measured).
	for (auto *J : B->jobs) tune_apply(J->plan, J->S, B->stream, true);
	// Layout: [jobs' data][output arena][work lists].
	size_t off = 0;
	for (auto *J : B->jobs) {
		J->off = off;
		off += J->S.bytes;
	}
	B->out_base = off;
	size_t arena = 0;
	for (auto *J : B->jobs) {
		J->out_off = arena;
		arena += (J->S.out_bytes + 255) & ~(size_t)255;
		ws_add(B->W, J->S, J->plan);
		B->out_pixels += (double)J->plan.out_width * J->plan.out_height;
		B->alg_bytes += (double)J->S.in_bytes + (double)J->S.out_bytes;
		B->flops += resample_flops(J->plan, J->S);
	}
	B->out_total = arena;
	ws_layout(B->W);
	B->block_bytes = off + arena + B->W.bytes;
	B->block = dev_alloc(B->block_bytes);
	if (!B->block) {
		for (auto *J : B->jobs) iw_set_error(J->ctx, "Out of device memory");
		B->failed = true;
		return 0;
	}
	char *base = (char *)B->block;
	bool ok = true;
	for (auto *J : B->jobs) {
		DevPlan &D = B->W.plans[J->index];
		fill_dev_plan(J->S, J->plan, base + J->off, D);
		D.out = (uint8_t *)(base + B->out_base + J->out_off);
		ok &= upload_job(J->S, J->plan, base + J->off, B->stream);
	}
	ok &= ws_upload(B->W, base + B->out_base + B->out_total, B->stream);
	ok &= hipStreamSynchronize(B->stream) == hipSuccess;
	// Host tables are no longer needed.
	for (auto *J : B->jobs) J->S = Staged();
	if (ok) {
		ok = hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
		if (ok) {
			ws_enqueue_resample(B->W, B->stream, true);
			ws_enqueue_errdiff(B->W, B->stream);
			ok = hipStreamEndCapture(B->stream, &B->graph) == hipSuccess;
		}
		ok = ok && hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0) == hipSuccess;
	}
	if (!ok) {
		for (auto *J : B->jobs) iw_set_error(J->ctx, "Batch upload failed");
		B->failed = true;
		return 0;
	}
	B->uploaded = true;
	return 1;
}

extern "C" int iwamd_batch_run(struct iwamd_batch *B)
{
	if (!B->uploaded && !iwamd_batch_upload(B)) return 0;
	hipEventRecord(B->ev[0], B->stream);
	const bool ok = hipGraphLaunch(B->exec, B->stream) == hipSuccess;
	hipEventRecord(B->ev[1], B->stream);
	if (!ok) {
		for (auto *J : B->jobs) iw_set_error(J->ctx, "Batch launch failed");
		return 0;
	}
	B->ran = true;
	return 1;
}

extern "C" int iwamd_batch_sync(struct iwamd_batch *B)
{
	const hipError_t e = hipStreamSynchronize(B->stream);
	if (e != hipSuccess) {
		for (auto *J : B->jobs) iw_set_errorf(J->ctx, "GPU batch failed: %s", hipGetErrorString(e));
		return 0;
	}
	return 1;
}

extern "C" int iwamd_batch_finish(struct iwamd_batch *B)
{
	if (!B->ran && !iwamd_batch_run(B)) return 0;
	std::vector<uint8_t> host(B->out_total + 1, 0);
	// The optimiser's scans of every output, one launch, before the copy.
	const size_t n = B->jobs.size();
	const size_t sbytes = opt_scan_bytes(n);
	std::vector<unsigned int> want(n, 0);
	std::vector<OptJob> scans;
	void *sblk = nullptr;
	for (size_t k = 0; k < n; k++)
		want[k] = B->jobs[k]->skip_pixels ? 0 : opt_scan_want(B->jobs[k]->ctx);
	bool ok = true;
	std::vector<iwx_opt_scan> res;
	if (std::any_of(want.begin(), want.end(), [](unsigned int w) { return w != 0; })) {
		sblk = dev_alloc(sbytes);
		ok = sblk != nullptr;
		if (ok) {
			iwx_opt_scan *dres = (iwx_opt_scan *)((char *)sblk + sbytes - n * sizeof(iwx_opt_scan));
			for (size_t k = 0; k < n; k++)
				if (want[k])
					scans.push_back(opt_job(B->jobs[k]->ctx, (char *)B->block + B->out_base + B->jobs[k]->out_off,
						dres + k, want[k]));
			res.resize(n);
			ok = opt_scan_enqueue(scans, sblk, n, B->stream) &&
				hipMemcpyAsync(res.data(), dres, n * sizeof(iwx_opt_scan), hipMemcpyDeviceToHost, B->stream) ==
					hipSuccess;
		}
	}
	ok = ok && hipMemcpyAsync(host.data(), (char *)B->block + B->out_base, B->out_total, hipMemcpyDeviceToHost,
		B->stream) == hipSuccess;
	ok = ok && iwamd_batch_sync(B);
	if (sblk) dev_release(sblk, sbytes);
	if (!ok) {
		for (auto *J : B->jobs) iw_set_error(J->ctx, "Device-to-host copy failed");
		return 0;
	}
	int all = 1;
	std::vector<uint8_t> zeros;
	for (size_t k = 0; k < n; k++) {
		auto *J = B->jobs[k];
		const uint8_t *img2 = host.data() + J->out_off;
		if (J->skip_pixels) {
			zeros.assign(J->plan.out_bpr * (size_t)J->plan.out_height + 1, 0);
			img2 = zeros.data();
		}
		const iwx_opt_scan *scan = nullptr;
		if (want[k]) {
			res[k].valid = want[k];
			scan = &res[k];
		}
		all &= iwx_finish_image(J->ctx, img2, scan);
	}
	return all;
}

extern "C" int iwamd_batch_kernel_ms(struct iwamd_batch *B, double *ms, int n)
{
	if (!B->ran || n < 1) return 0;
	float t = 0.0f;
	hipEventElapsedTime(&t, B->ev[0], B->ev[1]);
	ms[0] = t;
	return 1;
}

extern "C" double iwamd_batch_flops(struct iwamd_batch *B)
{
	return B->flops;
}

extern "C" void iwamd_batch_sizes(struct iwamd_batch *B, double *out_pixels, double *alg_bytes, int *n_jobs)
{
	if (out_pixels) *out_pixels = B->out_pixels;
	if (alg_bytes) *alg_bytes = B->alg_bytes;
	if (n_jobs) *n_jobs = (int)B->jobs.size();
}

extern "C" void iwamd_batch_destroy(struct iwamd_batch *B)
{
	if (!B) return;
	if (B->stream) hipStreamSynchronize(B->stream);
	if (B->exec) hipGraphExecDestroy(B->exec);
	if (B->graph) hipGraphDestroy(B->graph);
	dev_release(B->block, B->block_bytes);
	for (auto &e : B->ev)
		if (e) hipEventDestroy(e);
	if (B->stream) hipStreamDestroy(B->stream);
	for (auto *J : B->jobs) delete J;
	delete B;
}

// ---- request loop (continuous batching) -------------------------------------
// iwamd_server_submit() runs a request's host decisions and table building on
// the caller's thread and queues it.  The loop (iw_server.h: one worker
// thread) admits queued requests while their device bytes fit the HBM budget
// and launches every request admitted at the same moment as one flight on a
// free stream (one device block, persistent band kernels, error-diffusion
// planes, outputs copied back on the same stream).  Up to `streams` flights
// run at once, so a request that arrives while others are on the GPU starts
// as soon as memory and a stream free up, and its work overlaps theirs.  A
// request with error diffusion completes on its own: the last of its planes
// copies its output to coherent pinned host memory and raises a flag there.
// The rest complete with their flight.  Each request's bytes depend only on
// its own plan.
//
// Pinned host buffers and the per-slot events are made once and reused:
// freeing pinned memory waits for every stream of the device, which would
// tie each flight's end to all the others'.

namespace {

struct SrvReq {
	iw_context *ctx = nullptr;
	iwamd_plan plan;
	Staged S;
	size_t off = 0, out_off = 0;
	int widx = 0;
	int k = 0;                  // its place in the flight
	unsigned int scan_want = 0; // optimiser scans done on the device
	bool skip = false;
};

struct SrvFlight {
	WorkSet W;
	void *block = nullptr;
	size_t bytes = 0, out_base = 0, out_total = 0;
	uint8_t *host = nullptr;  // pinned, coherent: outputs, one flag per request, scan results
	size_t host_bytes = 0;
	int *flags = nullptr;
	size_t scan_dev = 0, scan_host = 0;  // scan region in the block; results in `host`
	std::vector<OptJob> scans;           // descriptors, kept until the flight ends
	hipEvent_t ev = nullptr;
};

// Pinned host buffers by size, reused across flights (worker thread only).
struct PinnedPool {
	std::multimap<size_t, uint8_t *> free;
	size_t cached = 0;
	static constexpr size_t kLimit = (size_t)4 << 30;
	uint8_t *get(size_t bytes, size_t &got)
	{
		auto it = free.lower_bound(bytes);
		if (it != free.end() && it->first <= 2 * bytes + (1 << 20)) {
			uint8_t *p = it->second;
			got = it->first;
			cached -= it->first;
			free.erase(it);
			return p;
		}
		uint8_t *p = nullptr;
		if (hipHostMalloc((void **)&p, bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return nullptr;
		got = bytes;
		return p;
	}
	void put(uint8_t *p, size_t bytes)
	{
		if (!p) return;
		if (cached + bytes > kLimit) {
			// The pool is full: this buffer goes back to the system.
			(void)hipHostFree(p);
			return;
		}
		free.emplace(bytes, p);
		cached += bytes;
	}
	void clear()
	{
		for (auto &kv : free) (void)hipHostFree(kv.second);
		free.clear();
		cached = 0;
	}
};

struct ServerExec {
	typedef SrvReq ReqData;
	typedef SrvFlight FlightData;
	typedef iwx::RequestLoop<ServerExec> Loop;
	int device = 0;
	std::vector<hipStream_t> streams;
	std::vector<hipEvent_t> events;  // one per stream slot (one flight per slot at a time)
	PinnedPool pool;

	void worker_start() { (void)hipSetDevice(device); }
	bool launch(Loop::Flight &F, int slot);
	bool flag_raised(Loop::Flight &F, int k)
	{
		return __atomic_load_n(&F.data.flags[k], __ATOMIC_ACQUIRE) == 1;
	}
	bool flight_done(Loop::Flight &F)
	{
		const hipError_t e = hipEventQuery(F.data.ev);
		if (e == hipErrorNotReady) return false;
		if (e != hipSuccess) F.failed = true;  // the flight cannot complete normally
		return true;
	}
	bool finish(Loop::Flight &F, Loop::Req &R, bool failed);
	void release(Loop::Flight &F);
};

// Bytes a request holds on the device while it runs.
size_t req_bytes(const Staged &S)
{
	return S.bytes + ((S.out_bytes + 255) & ~(size_t)255) + 64 * 1024;
}

bool ServerExec::launch(Loop::Flight &F, int slot)
{
	SrvFlight &D = F.data;
	const hipStream_t st = streams[slot];
	D.ev = events[slot];
	size_t off = 0;
	for (auto *R : F.reqs) {
		R->data.off = off;
		off += R->data.S.bytes;
	}
	D.out_base = off;
	size_t arena = 0;
	for (auto *R : F.reqs) {
		R->data.out_off = arena;
		arena += (R->data.S.out_bytes + 255) & ~(size_t)255;
		tune_apply(R->data.plan, R->data.S, st, false);
		R->data.widx = ws_add(D.W, R->data.S, R->data.plan);
	}
	D.out_total = arena;
	ws_layout(D.W);
	const size_t n = F.reqs.size();
	D.scan_dev = (off + arena + D.W.bytes + 255) & ~(size_t)255;
	D.bytes = D.scan_dev + opt_scan_bytes(n);
	D.block = dev_alloc(D.bytes);
	D.scan_host = (arena + sizeof(int) * n + 255) & ~(size_t)255;
	const size_t host_bytes = D.scan_host + n * sizeof(iwx_opt_scan);
	bool ok = D.block != nullptr;
	if (ok) {
		D.host = pool.get(host_bytes, D.host_bytes);
		ok = D.host != nullptr;
	}
	uint8_t *dhost = nullptr;
	if (ok) {
		D.flags = (int *)(D.host + arena);
		for (size_t k = 0; k < F.reqs.size(); k++) __atomic_store_n(&D.flags[k], 0, __ATOMIC_RELEASE);
		ok = hipHostGetDevicePointer((void **)&dhost, D.host, 0) == hipSuccess;
	}
	if (ok) {
		char *base = (char *)D.block;
		for (size_t k = 0; k < F.reqs.size(); k++) {
			SrvReq &R = F.reqs[k]->data;
			R.k = (int)k;
			R.scan_want = 0;
			DevPlan &P = D.W.plans[R.widx];
			fill_dev_plan(R.S, R.plan, base + R.off, P);
			P.out = (uint8_t *)(base + D.out_base + R.out_off);
			if (R.S.n_ed > 0 && !R.skip) {
				// The request completes on its own when its last plane ends.
				P.ed_done = (int *)(base + D.out_base + D.out_total + D.W.o_done) + R.widx;
				P.host_out = dhost + R.out_off;
				P.host_flag = (int *)(dhost + arena) + k;
				P.out_bytes = R.S.out_bytes;
				F.reqs[k]->flag = (int)k;
			}
			ok &= upload_job(R.S, R.plan, base + R.off, st);
		}
		ok &= ws_upload(D.W, base + D.out_base + D.out_total, st);
	}
	if (ok) {
		ws_enqueue_resample(D.W, st, true);
		ws_enqueue_errdiff(D.W, st);
		// The optimiser's scans of the outputs that complete with the
		// flight (one that completes on its own has been copied already).
		char *sreg = (char *)D.block + D.scan_dev;
		iwx_opt_scan *dres = (iwx_opt_scan *)(sreg + opt_scan_bytes(n) - n * sizeof(iwx_opt_scan));
		D.scans.clear();
		for (auto *R : F.reqs) {
			if (R->flag >= 0 || R->data.skip) continue;
			R->data.scan_want = opt_scan_want(R->data.ctx);
			if (R->data.scan_want)
				D.scans.push_back(opt_job(R->data.ctx, (char *)D.block + D.out_base + R->data.out_off,
					dres + R->data.k, R->data.scan_want));
		}
		if (!D.scans.empty()) {
			ok &= opt_scan_enqueue(D.scans, sreg, n, st);
			ok &= hipMemcpyAsync(D.host + D.scan_host, dres, n * sizeof(iwx_opt_scan), hipMemcpyDeviceToHost, st) ==
				hipSuccess;
		}
		// Requests without error diffusion complete with the flight.
		for (auto *R : F.reqs)
			if (R->flag < 0 && !R->data.skip)
				ok &= hipMemcpyAsync(D.host + R->data.out_off, (char *)D.block + D.out_base + R->data.out_off,
					R->data.S.out_bytes, hipMemcpyDeviceToHost, st) == hipSuccess;
	}
	// The event always ends the flight, also after a failed set-up.
	if (hipEventRecord(D.ev, st) != hipSuccess) ok = false;
	return ok;
}

// Completes one request from the flight's host memory.
bool ServerExec::finish(Loop::Flight &F, Loop::Req &R, bool failed)
{
	if (failed) {
		iw_set_error(R.data.ctx, "Request could not be run on the GPU");
		return false;
	}
	std::vector<uint8_t> zeros;
	const uint8_t *img2 = F.data.host + R.data.out_off;
	if (R.data.skip) {
		zeros.assign(R.data.plan.out_bpr * (size_t)R.data.plan.out_height + 1, 0);
		img2 = zeros.data();
	}
	std::unique_ptr<iwx_opt_scan> scan;
	if (R.data.scan_want) {
		scan.reset(new iwx_opt_scan());
		memcpy(scan.get(), F.data.host + F.data.scan_host + (size_t)R.data.k * sizeof(iwx_opt_scan),
			sizeof(iwx_opt_scan));
		scan->valid = R.data.scan_want;
	}
	return iwx_finish_image(R.data.ctx, img2, scan.get()) != 0;
}

void ServerExec::release(Loop::Flight &F)
{
	SrvFlight &D = F.data;
	// flight_done() saw the event complete: the stream is idle.
	(void)hipEventSynchronize(D.ev);
	for (auto *R : F.reqs) R->data.S = Staged();
	dev_release(D.block, D.bytes);
	pool.put(D.host, D.host_bytes);
	D.host = nullptr;
	D.block = nullptr;
}

}  // namespace

struct iwamd_server {
	ServerExec x;
	ServerExec::Loop *loop = nullptr;
	double timeout_ms = 300000.0;
	std::mutex err_mu;
	std::string error;
	void set_error(const std::string &s)
	{
		std::lock_guard<std::mutex> lock(err_mu);
		error = s;
		fprintf(stderr, "iwamd request loop: %s\n", s.c_str());
	}
};

extern "C" struct iwamd_server *iwamd_server_create(size_t hbm_budget_bytes, int streams)
{
	iwamd_server *V = new iwamd_server();
	if (streams < 1) streams = 1;
	if (streams > 16) streams = 16;
	if (hipGetDevice(&V->x.device) != hipSuccess) V->x.device = 0;
	for (int k = 0; k < streams; k++) {
		hipStream_t st = nullptr;
		hipEvent_t ev = nullptr;
		if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) break;
		if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
			hipStreamDestroy(st);
			break;
		}
		V->x.streams.push_back(st);
		V->x.events.push_back(ev);
	}
	if (V->x.streams.empty()) {
		delete V;
		return nullptr;
	}
	V->loop = new ServerExec::Loop(V->x, hbm_budget_bytes ? hbm_budget_bytes : ((size_t)64 << 30),
		(int)V->x.streams.size());
	V->loop->start();
	return V;
}

extern "C" int iwamd_server_submit(struct iwamd_server *V, struct iw_context *ctx)
{
	ServerExec::Loop::Req *R = new ServerExec::Loop::Req();
	R->data.ctx = ctx;
	if (!iwamd_prepare(ctx, &R->data.plan) || !runtime_init(ctx) ||
		!stage_job(ctx, R->data.plan, R->data.S, iw_get_option(ctx, "amd:resample"), true)) {
		delete R;
		return -1;
	}
	R->data.skip = R->data.S.skip_pixels;
	R->bytes = req_bytes(R->data.S);
	return V->loop->submit(R);
}

extern "C" void iwamd_server_set_timeout(struct iwamd_server *V, double ms)
{
	V->timeout_ms = ms;
}

extern "C" int iwamd_server_wait(struct iwamd_server *V, int id)
{
	std::string why;
	const int r = V->loop->wait(id, V->timeout_ms, &why);
	if (r < 0) {
		V->set_error(why);
		return 0;
	}
	return r;
}

extern "C" int iwamd_server_drain(struct iwamd_server *V)
{
	std::string why;
	const int r = V->loop->drain(V->timeout_ms, &why);
	if (r < 0) {
		V->set_error(why);
		return 0;
	}
	return r;
}

extern "C" const char *iwamd_server_error(struct iwamd_server *V)
{
	std::lock_guard<std::mutex> lock(V->err_mu);
	return V->error.empty() ? nullptr : V->error.c_str();
}

extern "C" int iwamd_server_request_ms(struct iwamd_server *V, int id, double *submit, double *start, double *done)
{
	return V->loop->times(id, submit, start, done) ? 1 : 0;
}

extern "C" void iwamd_server_stats(struct iwamd_server *V, size_t *bytes_held, size_t *max_bytes_held,
	int *flights)
{
	if (bytes_held) *bytes_held = V->loop->used();
	if (max_bytes_held) *max_bytes_held = V->loop->max_used();
	if (flights) *flights = V->loop->flights_launched();
}

extern "C" void iwamd_server_destroy(struct iwamd_server *V)
{
	if (!V) return;
	V->loop->stop();
	V->loop->delete_requests();
	delete V->loop;
	V->x.pool.clear();
	for (hipEvent_t ev : V->x.events) hipEventDestroy(ev);
	for (hipStream_t st : V->x.streams) hipStreamDestroy(st);
	delete V;
}


int iwx_gpu_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) return 0;
	return n;
}

// iw_process_image's pixel work: one job, run once.
int iwx_gpu_process(iw_context *ctx)
{
	iwamd_job *J = iwamd_job_create(ctx);
	if (!J) return 0;
	int ok = iwamd_job_run(J) && iwamd_job_finish(J);
	iwamd_job_destroy(J);
	return ok;
}
