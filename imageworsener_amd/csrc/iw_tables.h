// Host-built tables that drive the device kernels.
//
// Everything that needs libm (pow/sin/cos/exp) is evaluated here, on the host,
// with the system libm -- the device performs only IEEE + - * /, compares and
// conversions, so its results carry the same bits as the reference's.
#pragma once

#include "iwamd.h"

#include <cstdint>
#include <string>
#include <vector>

namespace iwx {

// One resampling pass along one dimension for one channel.  In WEIGHTS mode
// output o is sum over t in [start[o], start[o+1]) of in[src[t]] * w[t],
// accumulated from 0.0 in list order; src -1 reads the virtual-pixel value.
// In COPY mode output o is in[copy_src[o]], or 0.0 when copy_src[o] < 0.
struct WeightList {
	enum { WEIGHTS = 0, COPY = 1 };
	int mode = WEIGHTS;
	int num_in = 0, num_out = 0;
	int max_taps = 0;
	std::vector<int> start;      // num_out + 1
	std::vector<int> src;
	std::vector<double> w;
	std::vector<int> copy_src;   // num_out
	std::vector<int> pos;        // input position of each entry (may lie outside 0..num_in-1)
	double edge_value = 0.0;     // value of virtual pixels
	bool ok = true;              // false: unknown family
};

// The same list by input position: output o reads positions
// first[o] .. first[o]+count[o]-1 with weights w[off[o] + k], zero where the
// list has no entry (a dropped zero weight: adding v*0 to an accumulator that
// is never -0.0 leaves its bits unchanged).  A position outside 0..num_in-1
// reads the edge value (TRANSPARENT) or the nearest sample (REPLICATE); the
// STANDARD policy never lists one.  In COPY mode count is 1 (weight 1) or 0.
struct DenseWeights {
	std::vector<int> first, count, off;
	std::vector<double> w;
	int max_count = 0;
};
DenseWeights dense_weights(const WeightList &wl);

// src/imagew-resize.c:412-556 and :240-344.
WeightList build_weight_list(const iwamd_resize &rs, int channeltype, double edge_value);

// Quantisation of one output channel: linear sample s in [0,1] -> output code.
//
// With color_count==0 the reference computes e = clamp(F(s)*M) with F the
// colour-space transfer (libm pow) and picks floor(e) or ceil(e)
// (src/imagew-main.c:570-629).  Here that is replaced by thresholds over s
// found by bisection on the bit patterns of doubles:
//   T[k-1] = min{ s : e(s) >= k }   (k = 1..P)
//   U[k]   = min{ s : e(s) >  k }   (k = 0..P-1), U[P] = +inf
// so floor(e(s)) = #{k : T[k-1] <= s} and "e is an integer" <=> s < U[floor].
// code[k] is the output code of level k (posterised levels map to
// floor(0.5000000001 + k*(M/P))), lin[k] its value in linear light.
struct QuantTable {
	int P = 0;                   // number of levels - 1
	int ix_max = 0;              // most entries in one bin of ix
	std::vector<double> T, U, lin, code;
	std::vector<double> nc;      // nearest-colour midpoints (255), when used
	// The no-dither quantiser in one search (src/imagew-main.c:669-736):
	// level k = #{T <= s}; an exact sample (s < U[k]) keeps code[k], another
	// rounds up iff (lin[k+1] - s) <= (s - lin[k]), which holds exactly for
	// s >= M[k] (round_up_point).  So the code is code[#{k : D[k] <= s}] with
	// D[k] = min(T[k], max(T[k-1], U[k], M[k])) -- each D[k] lies in
	// [T[k-1], T[k]], so D is ascending and counts only its own level's step.
	std::vector<double> D;       // when folded (P entries)
	bool code_is_level = false;  // code[k] == k for every k
	// Bin index over the array the quantiser searches (nc when used, else T),
	// 2^ix_log2 bins of [0, 1]: ix[b] = #{m : a[m] < b / 2^ix_log2} for
	// b < 2^ix_log2, and the array's length last.  A sample s in bin b counts
	// ix[b] entries below it for certain and needs only a[ix[b] .. ix[b+1])
	// compared.
	std::vector<uint16_t> ix;
	int ix_log2 = 0;
};

// want_nc: nearest-colour midpoints.  fold_none (no dither, no nc): the
// searched array is QuantTable::D below instead of T.
QuantTable build_quant_table(const iwamd_ochan &oc, bool want_nc, bool fold_none = false);

// Decision data for error diffusion over a QuantTable (src/imagew-main.c:
// 730-748).  A non-exact sample s at level k is rounded up iff
// (lin[k+1] - s) <= (s - lin[k]) in FP64; both differences are monotone in s,
// so that holds exactly for s >= M[k] (M[P] = 2: never).
struct EdTables {
	std::vector<double> M;
};

EdTables build_ed_tables(const QuantTable &q);

// The smallest s in [0, 1] with (hi - s) <= (s - lo) in FP64 (2: none; the
// predicate is monotone in s).
double round_up_point(double lo, double hi);

// (T, U) for a transfer function and level count; cached per process.
const std::pair<std::vector<double>, std::vector<double>> &thresholds(const struct iw_csdescr &cs, int P);

// Linearising lookup for integer input samples: lut[v] = X(v / maxcc).
std::vector<double> build_input_lut(int nvalues, double maxcc, const struct iw_csdescr &cs);

// Random dithering: the reference's lag-1 complementary multiply-with-carry
// generator (src/imagew-util.c:199-243) viewed as Q_n = Q_1 * b^-(n-1) mod m,
// m = a*2^32 + 1, Q = a*x + c + 1.  row_q[j] = Q at the first draw of row j;
// col_mont[i] = b^-i * 2^64 mod m, so a Montgomery product gives the draw of
// pixel (i, j).
struct PrngTables {
	bool ok = true;
	std::vector<uint64_t> row_q;
	std::vector<uint64_t> col_mont;
};
PrngTables build_prng_tables(int seed, int width, int height);

constexpr uint64_t kPrngA = 0xfff0bf23ull;
constexpr uint64_t kPrngM = 0xfff0bf2300000001ull;
uint64_t prng_mont_minv();  // -m^-1 mod 2^64

// Reference-order CPU draws, for tests of the jump-ahead.
uint32_t prng_sequential(int seed, uint64_t n);

}  // namespace iwx
