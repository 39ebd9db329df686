// Host-built device tables: weight lists, linearising LUTs, quantisation
// thresholds, nearest-colour midpoints and PRNG jump-ahead constants.
#include "iw_tables.h"
#include "iw_ctx.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

namespace iwx {

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---- filter kernels (src/imagew-resize.c:62-185) --------------------------
// Each is evaluated at x >= 0; expressions keep the reference's operation
// order because the weights must carry the same bits.

struct FilterParams {
	int family;
	double radius, cubic_b, cubic_c, mix_param;
};

double sinc(double x)
{
	if (x <= 0.000000005) return 1.0;
	return sin(kPi * x) / (kPi * x);
}

double filter_eval(const FilterParams &f, double x)
{
	switch (f.family) {
	case IW_RESIZETYPE_LANCZOS:
		return x < f.radius ? sinc(x) * sinc(x / f.radius) : 0.0;
	case IW_RESIZETYPE_HANNING:
		return x < f.radius ? sinc(x) * (0.5 * cos(kPi * x / f.radius) + 0.5) : 0.0;
	case IW_RESIZETYPE_BLACKMAN:
		if (x < f.radius)
			return sinc(x) * (0.5 * cos(kPi * x / f.radius) + 0.08 * cos(2.0 * kPi * x / f.radius) + 0.42);
		return 0.0;
	case IW_RESIZETYPE_SINC:
		return x < f.radius ? sinc(x) : 0.0;
	case IW_RESIZETYPE_GAUSSIAN: {
		// Evaluated to 2.0 (4 sigma); tapered to zero over the last 0.001.
		if (x >= 2.0) return 0.0;
		double v = exp(-2.0 * x * x) * 0.79788456080286535587989;
		if (x <= 1.999) return v;
		return 1000.0 * (2.0 - x) * v;
	}
	case IW_RESIZETYPE_TRIANGLE:
		return x < 1.0 ? 1.0 - x : 0.0;
	case IW_RESIZETYPE_QUADRATIC:
		if (x < 0.5) return 0.75 - x * x;
		if (x < 1.5) return 0.50 * (x - 1.5) * (x - 1.5);
		return 0.0;
	case IW_RESIZETYPE_CUBIC: {
		// Mitchell-Netravali family with parameters B, C.
		const double b = f.cubic_b, c = f.cubic_c;
		if (x < 1.0)
			return ((12.0 - 9.0 * b - 6.0 * c) * x * x * x + (-18.0 + 12.0 * b + 6.0 * c) * x * x +
				(6.0 - 2.0 * b)) / 6.0;
		if (x < 2.0)
			return ((-b - 6.0 * c) * x * x * x + (6.0 * b + 30.0 * c) * x * x +
				(-12.0 * b - 48.0 * c) * x + (8.0 * b + 24.0 * c)) / 6.0;
		return 0.0;
	}
	case IW_RESIZETYPE_HERMITE:
		return x < 1.0 ? 2.0 * x * x * x - 3.0 * x * x + 1.0 : 0.0;
	case IW_RESIZETYPE_BOX:
		return x <= 0.5 ? 1.0 : 0.0;
	case IW_RESIZETYPE_BOXAVG:
		if (x < 0.4999999) return 1.0;
		if (x <= 0.5000001) return 0.5;
		return 0.0;
	case IW_RESIZETYPE_MIX:
		if (x < (0.5 - f.mix_param / 2.0)) return 1.0;
		if (x < (0.5 + f.mix_param / 2.0)) return 0.5 - (x - 0.5) / f.mix_param;
		return 0.0;
	}
	return 0.0;
}

}  // namespace

WeightList build_weight_list(const iwamd_resize &rs, int channeltype, double edge_value)
{
	WeightList wl;
	wl.num_in = rs.num_in;
	wl.num_out = rs.num_out;
	wl.edge_value = edge_value;

	FilterParams f = {rs.family, 0.0, 0.0, 0.0, 0.0};
	bool boxhack = false, standard = true;
	switch (rs.family) {
	case IW_RESIZETYPE_NULL: standard = false; break;
	case IW_RESIZETYPE_NEAREST: standard = false; boxhack = true; break;
	case IW_RESIZETYPE_MIX:
		// Trapezoid whose shape follows the scale factor.
		f.mix_param = ((double)rs.num_out) / rs.num_in;
		if (f.mix_param > 1.0) f.mix_param = 1.0 / f.mix_param;
		f.radius = 0.5 + f.mix_param;
		break;
	case IW_RESIZETYPE_BOX: f.radius = 0.501; boxhack = true; break;
	case IW_RESIZETYPE_BOXAVG: f.radius = 0.501; break;
	case IW_RESIZETYPE_TRIANGLE: f.radius = 1.0; break;
	case IW_RESIZETYPE_QUADRATIC: f.radius = 1.5; break;
	case IW_RESIZETYPE_GAUSSIAN: f.radius = 2.0; break;
	case IW_RESIZETYPE_HERMITE: f.radius = 1.0; break;
	case IW_RESIZETYPE_CUBIC:
		f.radius = 2.0;
		f.cubic_b = rs.param1 < -100.0 ? -100.0 : (rs.param1 > 100.0 ? 100.0 : rs.param1);
		f.cubic_c = rs.param2 < -100.0 ? -100.0 : (rs.param2 > 100.0 ? 100.0 : rs.param2);
		break;
	case IW_RESIZETYPE_LANCZOS:
	case IW_RESIZETYPE_HANNING:
	case IW_RESIZETYPE_BLACKMAN:
	case IW_RESIZETYPE_SINC:
		// Radius is the number of lobes, 2..100.
		f.radius = floor(rs.param1 + 0.5);
		if (f.radius < 2.0) f.radius = 2.0;
		if (f.radius > 100.0) f.radius = 100.0;
		break;
	default:
		wl.ok = false;
		return wl;
	}

	double blur = rs.blur_factor;
	if (blur < 0.0001) blur = 0.0001;
	if (blur > 10000.0) blur = 10000.0;
	double offset = rs.translate;
	// Shift box-like filters a hair so no pixel sits on a box border.
	if (boxhack) offset -= 0.00000000001;
	if (rs.use_offset && channeltype >= 0 && channeltype <= 2) offset += rs.channel_offset[channeltype];

	if (rs.family == IW_RESIZETYPE_NULL) {
		wl.mode = WeightList::COPY;
		wl.copy_src.resize(rs.num_out);
		for (int o = 0; o < rs.num_out; o++) wl.copy_src[o] = (o < rs.num_in) ? o : -1;
		wl.max_taps = 1;
		return wl;
	}
	if (rs.family == IW_RESIZETYPE_NEAREST) {
		// Literal nearest neighbour on the output grid, ignoring blur
		// (src/imagew-resize.c:377-393).
		wl.mode = WeightList::COPY;
		wl.copy_src.resize(rs.num_out);
		for (int o = 0; o < rs.num_out; o++) {
			double center = (0.5 + (double)o - offset) / (double)rs.num_out;
			int ip = (int)floor(center * (double)rs.num_in);
			if (ip < 0) ip = 0;
			else if (ip > rs.num_in - 1) ip = rs.num_in - 1;
			wl.copy_src[o] = ip;
		}
		wl.max_taps = 1;
		return wl;
	}
	(void)standard;

	double reduction = (rs.out_true_size < (double)rs.num_in) ? ((double)rs.num_in) / rs.out_true_size : 1.0;
	reduction *= blur;

	wl.start.resize(rs.num_out + 1);
	for (int o = 0; o < rs.num_out; o++) {
		wl.start[o] = (int)wl.src.size();
		double center = (0.5 + (double)o - offset) / rs.out_true_size;
		double pos_in = center * (double)rs.num_in - 0.5;
		int first = (int)ceil(pos_in - f.radius * reduction - 0.0001);
		int last = (int)floor(pos_in + f.radius * reduction + 0.0001);
		double v_sum = 0.0;
		int v_count = 0;
		const size_t begin = wl.src.size();
		for (int i = first; i <= last; i++) {
			const bool outside = (i < 0 || i >= rs.num_in);
			if (rs.edge_policy == IW_EDGE_POLICY_STANDARD && outside) continue;
			double pos = (((double)i) - pos_in) / reduction;
			double v = filter_eval(f, pos < 0 ? -pos : pos);
			if (v == 0.0) continue;
			v_sum += v;
			v_count++;
			int s;
			if (!outside) s = i;
			else if (rs.edge_policy == IW_EDGE_POLICY_TRANSPARENT) s = -1;
			else s = (i < 0) ? 0 : rs.num_in - 1;
			wl.src.push_back(s);
			wl.pos.push_back(i);
			wl.w.push_back(v);
		}
		// Normalise by division by the running sum, in list order; a zero sum
		// zeroes the weights instead.
		if (v_count > 0) {
			for (size_t t = begin; t < wl.src.size(); t++) wl.w[t] = (v_sum != 0.0) ? wl.w[t] / v_sum : 0.0;
		}
		int n = (int)(wl.src.size() - begin);
		if (n > wl.max_taps) wl.max_taps = n;
	}
	wl.start[rs.num_out] = (int)wl.src.size();
	return wl;
}

// ---- quantisation tables --------------------------------------------------

namespace {

uint64_t dbits(double d)
{
	uint64_t u;
	memcpy(&u, &d, 8);
	return u;
}

double bitsd(uint64_t u)
{
	double d;
	memcpy(&d, &u, 8);
	return d;
}

// e(s) of get_nearest_valid_colors for levels P (src/imagew-main.c:581-607).
double expanded(double s, const struct iw_csdescr &cs, double P)
{
	double e = iwx_linear_to_x(s, &cs) * P;
	if (e > P) e = P;
	if (e < 0.0) e = 0.0;
	return e;
}

// Smallest double s in [0,1] with pred(s); pred(1.0) must hold and pred must
// be monotone.  Non-negative doubles order like their bit patterns.
template <class Pred>
double lowest_true(Pred pred)
{
	uint64_t lo = 0, hi = dbits(1.0);
	if (pred(0.0)) return 0.0;
	while (hi - lo > 1) {
		uint64_t mid = lo + (hi - lo) / 2;
		if (pred(bitsd(mid))) hi = mid;
		else lo = mid;
	}
	return bitsd(hi);
}

struct ThreshKey {
	int cstype;
	double gamma;
	int P;
	bool operator<(const ThreshKey &o) const
	{
		return std::tie(cstype, gamma, P) < std::tie(o.cstype, o.gamma, o.P);
	}
};

std::mutex g_thresh_mutex;
std::map<ThreshKey, std::pair<std::vector<double>, std::vector<double>>> g_thresh_cache;

// Thresholds depend only on the transfer function and P; cached per process
// (a 16-bit table costs ~4M libm calls).
}  // namespace

const std::pair<std::vector<double>, std::vector<double>> &thresholds(const struct iw_csdescr &cs, int P)
{
	ThreshKey key = {cs.cstype, cs.cstype == IW_CSTYPE_GAMMA ? cs.gamma : 0.0, P};
	std::lock_guard<std::mutex> lock(g_thresh_mutex);
	auto it = g_thresh_cache.find(key);
	if (it != g_thresh_cache.end()) return it->second;
	std::vector<double> T(P), U(P + 1);
	const double Pd = (double)P;
	for (int k = 1; k <= P; k++) {
		const double kd = (double)k;
		T[k - 1] = lowest_true([&](double s) { return expanded(s, cs, Pd) >= kd; });
	}
	for (int k = 0; k < P; k++) {
		const double kd = (double)k;
		U[k] = lowest_true([&](double s) { return expanded(s, cs, Pd) > kd; });
	}
	U[P] = INFINITY;
	return g_thresh_cache.emplace(key, std::make_pair(std::move(T), std::move(U))).first->second;
}

double round_up_point(double lo, double hi)
{
	auto ceil_wins = [&](double s) { return (hi - s) <= (s - lo); };
	if (ceil_wins(0.0)) return 0.0;
	if (!ceil_wins(1.0)) return 2.0;
	uint64_t a = 0, b = dbits(1.0);  // !ceil_wins(a), ceil_wins(b)
	while (b - a > 1) {
		const uint64_t mid = a + (b - a) / 2;
		if (ceil_wins(bitsd(mid))) b = mid;
		else a = mid;
	}
	return bitsd(b);
}

namespace {
struct QuantKey {
	int cstype;
	double gamma, maxcc;
	int P, color_count, nc, fold;
	bool operator<(const QuantKey &o) const
	{
		return std::tie(cstype, gamma, maxcc, P, color_count, nc, fold) <
		       std::tie(o.cstype, o.gamma, o.maxcc, o.P, o.color_count, o.nc, o.fold);
	}
};
std::mutex g_quant_mutex;
std::map<QuantKey, QuantTable> g_quant_cache;
QuantTable build_quant_table_uncached(const iwamd_ochan &oc, bool want_nc, bool fold_none);
}  // namespace

// A table depends only on the output transfer function, the code range and
// the options; cached per process (most jobs of a batch share one).
QuantTable build_quant_table(const iwamd_ochan &oc, bool want_nc, bool fold_none)
{
	const QuantKey key = {oc.out_cs.cstype, oc.out_cs.cstype == IW_CSTYPE_GAMMA ? oc.out_cs.gamma : 0.0,
		oc.maxcolorcode_dbl, oc.maxcolorcode_int, oc.color_count, want_nc ? 1 : 0, fold_none ? 1 : 0};
	{
		std::lock_guard<std::mutex> lock(g_quant_mutex);
		auto it = g_quant_cache.find(key);
		if (it != g_quant_cache.end()) return it->second;
	}
	QuantTable q = build_quant_table_uncached(oc, want_nc, fold_none);
	std::lock_guard<std::mutex> lock(g_quant_mutex);
	return g_quant_cache.emplace(key, std::move(q)).first->second;
}

namespace {
QuantTable build_quant_table_uncached(const iwamd_ochan &oc, bool want_nc, bool fold_none)
{
	QuantTable q;
	const double M = oc.maxcolorcode_dbl;
	const int P = oc.color_count ? oc.color_count - 1 : oc.maxcolorcode_int;
	q.P = P;
	const auto &tu = thresholds(oc.out_cs, P);
	q.T = tu.first;
	q.U = tu.second;
	q.code.resize(P + 1);
	q.lin.resize(P + 1);
	for (int k = 0; k <= P; k++) {
		double code = (double)k;
		if (oc.color_count) code = floor(0.5000000001 + ((double)k) * (M / (double)P));
		q.code[k] = code;
		// Candidate back in linear light (src/imagew-main.c:330-344).
		unsigned int ci = (unsigned int)code;
		q.lin[k] = (oc.out_cs.cstype == IW_CSTYPE_LINEAR) ? ((double)ci) / M
			: iwx_x_to_linear(((double)ci) / M, &oc.out_cs);
	}
	if (want_nc) {
		// Midpoints in linear light between adjacent codes
		// (src/imagew-main.c:1197-1236).
		q.nc.resize(255);
		double prev = 0.0;
		for (int i = 0; i < 255; i++) {
			double cur = iwx_x_to_linear(((double)(i + 1)) / 255, &oc.out_cs);
			q.nc[i] = (prev + cur) / 2.0;
			prev = cur;
		}
	}
	q.code_is_level = true;
	for (int k = 0; k <= P; k++)
		if (q.code[k] != (double)k) q.code_is_level = false;
	if (fold_none && !want_nc) {
		q.D.resize(P);
		for (int k = 0; k < P; k++) {
			const double m = round_up_point(q.lin[k], q.lin[k + 1]);
			double d = std::max(q.U[k], m);
			if (k > 0) d = std::max(d, q.T[k - 1]);
			q.D[k] = std::min(q.T[k], d);
		}
	}
	// Bins: about eight per searched entry, 2^11..2^16 of them.
	const std::vector<double> &a = want_nc ? q.nc : (fold_none ? q.D : q.T);
	const int n = want_nc ? (int)q.nc.size() : P;
	if (n > 65535) return q;  // counts would not fit: no index, full search
	int L = 11;
	while (L < 16 && (1 << L) < 8 * n) L++;
	const int NB = 1 << L;
	q.ix_log2 = L;
	q.ix.resize(NB + 1);
	int m = 0;
	for (int b = 0; b < NB; b++) {
		const double edge = (double)b / (double)NB;
		while (m < n && a[m] < edge) m++;
		q.ix[b] = (uint16_t)m;
	}
	q.ix[NB] = (uint16_t)n;
	for (int b = 0; b < NB; b++) q.ix_max = std::max(q.ix_max, (int)(q.ix[b + 1] - q.ix[b]));
	return q;
}
}  // namespace

DenseWeights dense_weights(const WeightList &wl)
{
	DenseWeights d;
	const int n = wl.num_out;
	d.first.assign(n, 0);
	d.count.assign(n, 0);
	d.off.assign(n, 0);
	for (int o = 0; o < n; o++) {
		d.off[o] = (int)d.w.size();
		if (wl.mode == WeightList::COPY) {
			if (wl.copy_src[o] >= 0) {
				d.first[o] = wl.copy_src[o];
				d.count[o] = 1;
				d.w.push_back(1.0);
			}
		} else if (wl.start[o + 1] > wl.start[o]) {
			const int a = wl.start[o], b = wl.start[o + 1];
			const int p0 = wl.pos[a], p1 = wl.pos[b - 1];
			d.first[o] = p0;
			d.count[o] = p1 - p0 + 1;
			d.w.resize(d.off[o] + d.count[o], 0.0);
			for (int t = a; t < b; t++) d.w[d.off[o] + (wl.pos[t] - p0)] = wl.w[t];
		}
		d.max_count = std::max(d.max_count, d.count[o]);
	}
	return d;
}

EdTables build_ed_tables(const QuantTable &q)
{
	EdTables t;
	const int P = q.P;
	t.M.assign(P + 1, 2.0);
	for (int k = 0; k < P; k++) t.M[k] = round_up_point(q.lin[k], q.lin[k + 1]);
	return t;
}

}  // namespace iwx

extern "C" int iwamd_quant_thresholds(int cstype, double gamma, int P, double *T, double *U)
{
	struct iw_csdescr cs = {cstype, 0, gamma};
	const auto &tu = iwx::thresholds(cs, P);
	for (int k = 0; k < P; k++) T[k] = tu.first[k];
	for (int k = 0; k <= P; k++) U[k] = tu.second[k];
	return 1;
}

namespace iwx {

namespace {
std::mutex g_lut_mutex;
std::map<std::tuple<int, int, double, double>, std::vector<double>> g_lut_cache;
}  // namespace

// Cached per process: a 16-bit table costs 65536 libm calls.
std::vector<double> build_input_lut(int nvalues, double maxcc, const struct iw_csdescr &cs)
{
	const auto key = std::make_tuple(nvalues, cs.cstype, cs.cstype == IW_CSTYPE_GAMMA ? cs.gamma : 0.0, maxcc);
	{
		std::lock_guard<std::mutex> lock(g_lut_mutex);
		auto it = g_lut_cache.find(key);
		if (it != g_lut_cache.end()) return it->second;
	}
	std::vector<double> lut(nvalues);
	for (int v = 0; v < nvalues; v++) {
		double s = ((double)v) / maxcc;
		lut[v] = (cs.cstype == IW_CSTYPE_LINEAR) ? s : iwx_x_to_linear(s, &cs);
	}
	std::lock_guard<std::mutex> lock(g_lut_mutex);
	return g_lut_cache.emplace(key, std::move(lut)).first->second;
}

// ---- PRNG jump-ahead ------------------------------------------------------

namespace {

typedef unsigned __int128 u128;

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)(((u128)a * b) % m); }

uint64_t powmod(uint64_t base, uint64_t e, uint64_t m)
{
	uint64_t r = 1 % m;
	base %= m;
	while (e) {
		if (e & 1) r = mulmod(r, base, m);
		base = mulmod(base, base, m);
		e >>= 1;
	}
	return r;
}

}  // namespace

uint64_t prng_mont_minv()
{
	// Newton iteration for m^-1 mod 2^64, negated.
	uint64_t inv = kPrngM;  // correct to 1 bit... m odd: inv*m == 1 mod 2^3
	for (int i = 0; i < 6; i++) inv *= 2 - kPrngM * inv;
	return (uint64_t)0 - inv;
}

// One draw of the reference generator (src/imagew-util.c:220-233).
static uint32_t prng_step(uint32_t &x, uint32_t &c)
{
	uint64_t t = kPrngA * (uint64_t)x + c;
	c = (uint32_t)(t >> 32);
	x = 0xffffffffu - (uint32_t)(t & 0xffffffffu);
	return x;
}

uint32_t prng_sequential(int seed, uint64_t n)
{
	uint32_t x = 0x03333333u + (uint32_t)seed, c = 0x05555555u + (uint32_t)seed, r = 0;
	for (uint64_t i = 0; i < n; i++) r = prng_step(x, c);
	return r;
}

PrngTables build_prng_tables(int seed, int width, int height)
{
	PrngTables t;
	// Take the first draw on the host; from there the state is canonical
	// (carry < a) and Q_n = Q_1 * b^-(n-1) mod m holds.
	uint32_t x = 0x03333333u + (uint32_t)seed, c = 0x05555555u + (uint32_t)seed;
	prng_step(x, c);
	if ((uint64_t)c >= kPrngA) {
		t.ok = false;
		return t;
	}
	const uint64_t m = kPrngM;
	const uint64_t q1 = kPrngA * (uint64_t)x + c + 1;
	const uint64_t binv = m - kPrngA;  // 2^32 * (-a) == 1 (mod m)
	const uint64_t row_step = powmod(binv, (uint64_t)width, m);
	t.row_q.resize(height);
	uint64_t q = q1;
	for (int j = 0; j < height; j++) {
		t.row_q[j] = q;
		q = mulmod(q, row_step, m);
	}
	// Montgomery form of b^-i: b^-i * 2^64 mod m.
	const uint64_t r64 = (uint64_t)(((u128)1 << 64) % m);
	t.col_mont.resize(width);
	uint64_t p = r64;
	for (int i = 0; i < width; i++) {
		t.col_mont[i] = p;
		p = mulmod(p, binv, m);
	}
	return t;
}

}  // namespace iwx

extern "C" int iwamd_ed_tables(const struct iwamd_ochan *och, double *lin, double *M)
{
	if (och->use_nc_table) return -1;
	const iwx::QuantTable q = iwx::build_quant_table(*och, false);
	const iwx::EdTables t = iwx::build_ed_tables(q);
	for (int k = 0; k <= q.P; k++) {
		lin[k] = q.lin[k];
		M[k] = t.M[k];
	}
	return q.P;
}
