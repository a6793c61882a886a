// encode_normal.h - MODE_NORMAL block encoder: all-16 and random candidates.
//
// Follows s2tc_encode_block<dxt, ColorDist, MODE_NORMAL, refine>
// (s2tc_algorithm.cpp:936-1107).  One block is handled by one DPP row of 16
// lanes; lane t owns texel t (x = t >> 2, y = t & 3, i.e. the reference's
// x-major visiting order).  A 64-lane workgroup holds four blocks.
//
// Work split:
//  * candidate gather: ballot + prefix count gives each opaque texel its
//    position in the reference's candidate list (:940-951);
//  * random candidates: lane t draws candidates t, t+16, ... with the
//    counter-based generator (s2tc_prng.h) in the reference's draw order;
//  * distance matrix: lane t computes rows i = 16p + t of dists[i][k]
//    (:373-392): the first two row blocks into its registers, the rest
//    into LDS;
//  * pair scan (reduce_colors_inplace, :393-410): for every row-block pair
//    (p, q) lane t keeps its row of block p and its row of block q in
//    registers and meets the 16 rows of block q by rotating them through the
//    DPP row with row_ror; each pair's sum is formed once, by one lane;
//  * argmin: the reference's fold "if (bestsum < 0 || sum < bestsum)" equals
//    the first strict minimum over the pairs that come after the last pair
//    with a negative (int32-wrapped, SRGB only) sum, or that last pair if no
//    pair follows it.  Lanes reduce (sum, pair order) keys with DPP; only a
//    block that has a negative sum runs the second pass.
//  * refinement: lane t scores its own texel, the cluster sums are row
//    reductions (4 DPP steps each).
#pragma once

#include "encode_fast.h"
#include "kernel_table.h"  // S2TC_REG_RB

namespace s2tc {

#define S2TC_RS 16  // LDS row stride in ints: no padding, so that 11 four-block workgroups of m = 80 fit a CU

// Every control used here (quad_perm, row mirrors, row_ror) reads a lane of
// the same row, and rows are all-active or all-inactive, so bound_ctrl never
// applies; the update_dpp form with bound_ctrl lets the compiler fold the
// lane move into the consuming VALU op (v_min_i32_dpp, v_add_u32_dpp).
template <int CTRL> S2TC_D int dpp_i(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true); }
template <int CTRL> S2TC_D uint32_t dpp_u(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true); }

// all-reduce within a DPP row of 16 lanes: quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror
#define S2TC_ROW_STEPS(OP) OP(0xB1) OP(0x4E) OP(0x141) OP(0x140)
S2TC_D uint32_t row_sum(uint32_t v)
{
#define S2TC_STEP(C) v += dpp_u<C>(v);
	S2TC_ROW_STEPS(S2TC_STEP)
#undef S2TC_STEP
	return v;
}
S2TC_D uint32_t row_or(uint32_t v)
{
#define S2TC_STEP(C) v |= dpp_u<C>(v);
	S2TC_ROW_STEPS(S2TC_STEP)
#undef S2TC_STEP
	return v;
}
S2TC_D int row_min(int v)
{
#define S2TC_STEP(C) { int o = dpp_i<C>(v); v = o < v ? o : v; }
	S2TC_ROW_STEPS(S2TC_STEP)
#undef S2TC_STEP
	return v;
}
S2TC_D int row_max(int v)
{
#define S2TC_STEP(C) { int o = dpp_i<C>(v); v = o > v ? o : v; }
	S2TC_ROW_STEPS(S2TC_STEP)
#undef S2TC_STEP
	return v;
}
// lexicographic (hi, lo) unsigned minimum
S2TC_D void row_min_key(uint32_t &hi, uint32_t &lo)
{
#define S2TC_STEP(C) { uint32_t oh = dpp_u<C>(hi), ol = dpp_u<C>(lo); \
	bool lt = oh < hi || (oh == hi && ol < lo); hi = lt ? oh : hi; lo = lt ? ol : lo; }
	S2TC_ROW_STEPS(S2TC_STEP)
#undef S2TC_STEP
}

struct ScanKey {
	uint32_t hi, lo;  // (sum, pair order i << 16 | j); all-ones = none
	int qneg;         // largest pair order with a negative sum, -1 if none
};

// One rotation step of the pair scan: pairs (16p + t, 16q + partner) where
// partner is the lane whose row of block q arrives via row_ror:R.
template <int R, bool DIAG, int PASS, bool NEG>
S2TC_D void scan_step(const int (&X)[16], const int (&Y)[16], int t, int p, int q, int m, int qthr, ScanKey &st)
{
	int s = 0;
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		int y = R ? dpp_i<0x120 + R>(Y[k]) : Y[k];
		s = (int)((uint32_t)s + (uint32_t)(X[k] < y ? X[k] : y));  // int32 wrap as in the reference
	}
	const int b = R ? dpp_i<0x120 + R>(t) : t;
	const int i = 16 * p + t, j = 16 * q + b;
	bool valid = i < m && j < m;
	if (DIAG && R == 8) valid = valid && t < 8;
	const int lo = i < j ? i : j, hi = i < j ? j : i;
	const uint32_t order = ((uint32_t)lo << 16) | (uint32_t)hi;
	if (PASS == 2) valid = valid && (int)order > qthr;
	const bool better = valid && ((uint32_t)s < st.hi || ((uint32_t)s == st.hi && order < st.lo));
	st.hi = better ? (uint32_t)s : st.hi;
	st.lo = better ? order : st.lo;
	if (NEG && PASS == 1) {
		const bool neg = valid && s < 0;
		st.qneg = (neg && (int)order > st.qneg) ? (int)order : st.qneg;
	}
}

template <int R, int END, bool DIAG, int PASS, bool NEG> struct ScanSteps {
	static S2TC_D void run(const int (&X)[16], const int (&Y)[16], int t, int p, int q, int m, int qthr, ScanKey &st)
	{
		scan_step<R, DIAG, PASS, NEG>(X, Y, t, p, q, m, qthr, st);
		ScanSteps<R + 1, END, DIAG, PASS, NEG>::run(X, Y, t, p, q, m, qthr, st);
	}
};
template <int END, bool DIAG, int PASS, bool NEG> struct ScanSteps<END, END, DIAG, PASS, NEG> {
	static S2TC_D void run(const int (&)[16], const int (&)[16], int, int, int, int, int, ScanKey &) {}
};

// The first S2TC_REG_RB row blocks of a block's distance matrix stay in the
// registers of their lanes (RegRows); the row memory holds the rows from
// 16 S2TC_REG_RB on.  Fewer rows in LDS let more workgroups share a CU.
struct RegRows {
	int r[S2TC_REG_RB][16];
};

// Row r >= 16 S2TC_REG_RB from the row memory.
S2TC_D void load_row(const int *rows, int r, int (&X)[16])
{
	const int4 *p = (const int4 *)(rows + (r - 16 * S2TC_REG_RB) * S2TC_RS);
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		int4 v = p[k];
		X[4 * k] = v.x;
		X[4 * k + 1] = v.y;
		X[4 * k + 2] = v.z;
		X[4 * k + 3] = v.w;
	}
}

// Lane t's row of row block p (row 16p + t), wherever it lives.
S2TC_D void get_rows(const int *rows, const RegRows &R, int p, int t, int (&X)[16])
{
	if (p >= S2TC_REG_RB) {
		load_row(rows, 16 * p + t, X);
		return;
	}
#pragma unroll
	for (int z = 0; z < S2TC_REG_RB; ++z)
		if (p == z) {
#pragma unroll
			for (int k = 0; k < 16; ++k) X[k] = R.r[z][k];
		}
}
S2TC_D void put_rows(int *rows, RegRows &R, int p, int t, const int (&dv)[16])
{
	if (p >= S2TC_REG_RB) {
		int4 *rp = (int4 *)(rows + (16 * p + t - 16 * S2TC_REG_RB) * S2TC_RS);
#pragma unroll
		for (int k = 0; k < 4; ++k) rp[k] = make_int4(dv[4 * k], dv[4 * k + 1], dv[4 * k + 2], dv[4 * k + 3]);
		return;
	}
#pragma unroll
	for (int z = 0; z < S2TC_REG_RB; ++z)
		if (p == z) {
#pragma unroll
			for (int k = 0; k < 16; ++k) R.r[z][k] = dv[k];
		}
}

// Scan over all pairs i < j < m of the rows in LDS; pbw row blocks (uniform).
template <int PASS, bool NEG>
S2TC_D void pair_scan(const int *rows, const RegRows &R, int t, int m, int pbw, int qthr, ScanKey &st)
{
	for (int p = 0; p < pbw; ++p) {
		int X[16];
		get_rows(rows, R, p, t, X);
		for (int q = p; q < pbw; ++q) {
			int Y[16];
			get_rows(rows, R, q, t, Y);
			if (q == p)
				ScanSteps<1, 9, true, PASS, NEG>::run(X, Y, t, p, q, m, qthr, st);
			else
				ScanSteps<0, 16, false, PASS, NEG>::run(X, Y, t, p, q, m, qthr, st);
		}
	}
}

// Merge a (sum, pair order) candidate into the running lexicographic minimum.
S2TC_D void merge_key(uint32_t s, uint32_t order, bool ok, ScanKey &st)
{
	const bool better = ok && (s < st.hi || (s == st.hi && order < st.lo));
	st.hi = better ? s : st.hi;
	st.lo = better ? order : st.lo;
}

// The pair scan for rows of non-negative values whose 16-term sums do not
// wrap (every distance but SRGB, and DXT5 alpha).  Within one row-block pair
// a lane folds its pairs into one key per step, (sum << 4 | b) for the
// partner b (lane within the partner's row block).  Off the diagonal the
// pair order is i << 16 | 16q + b; on it, the pairs (t, b) of lane t come in
// the order (min, max), which grows with b (partners below t first, each
// meeting t as the first of the pair).  So the smallest key is the smallest
// (sum, order).  The partner's index travels through the DPP row with its
// row values, all-ones when its row is outside the candidates, which makes
// the key of an invalid pair all-ones, larger than every valid one.  Key
// kinds:
//   KEY_NARROW  32 bits, sums below 2^23 (bounded distances, alpha);
//   KEY_SAT     32 bits, the sum saturated at KEY_SAT_MAX: exact wherever the
//               block's least sum lies below it, which the caller checks
//               (a block at the cap scans again with KEY_WIDE);
//   KEY_WIDE    64 bits, sums below 2^31.
// The running (sum, order) minimum is merged once per row-block pair.
enum { KEY_NARROW = 0, KEY_SAT = 1, KEY_WIDE = 2 };
#define KEY_SAT_MAX 0x0ffffffeu  // (KEY_SAT_MAX << 4 | 15) stays below all-ones
template <int KM> struct ScanKeyT { typedef uint32_t T; };
template <> struct ScanKeyT<KEY_WIDE> { typedef uint64_t T; };

// The cap of the KEY_SAT sums: KEY_SAT_MAX, or lower for tests (set by
// k_set_scan_cap), which sends blocks through the exact rescan.
static __device__ uint32_t s2tc_scan_cap = KEY_SAT_MAX;

template <int KM> S2TC_D typename ScanKeyT<KM>::T pair_key(int s, uint32_t b, uint32_t cap)
{
	if constexpr (KM == KEY_WIDE) {
		// an invalid partner (all-ones) sets the high word to >= 2^28 - 1 > s >> 28
		return (((uint64_t)(uint32_t)s << 4) | b) | ((uint64_t)(b >> 4) << 32);
	} else if constexpr (KM == KEY_SAT) {
		return (min((uint32_t)s, cap) << 4) | b;
	} else {
		return ((uint32_t)s << 4) | b;
	}
}
template <int KM> S2TC_D bool key_valid(typename ScanKeyT<KM>::T k)
{
	if constexpr (KM == KEY_WIDE) return (k >> 32) < 8u;
	else return k != ~0u;
}

template <int R, int END, int KM> struct KeyStepsOff {
	typedef typename ScanKeyT<KM>::T K;
	static S2TC_D void run(const int (&X)[16], const int (&Y)[16], uint32_t w, uint32_t cap, K &kmin)
	{
		int s = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const int y = R ? dpp_i<0x120 + R>(Y[k]) : Y[k];
			s += X[k] < y ? X[k] : y;
		}
		const uint32_t b = R ? dpp_u<0x120 + R>(w) : w;
		kmin = min(kmin, pair_key<KM>(s, b, cap));
		KeyStepsOff<R + 1, END, KM>::run(X, Y, w, cap, kmin);
	}
};
template <int END, int KM> struct KeyStepsOff<END, END, KM> {
	static S2TC_D void run(const int (&)[16], const int (&)[16], uint32_t, uint32_t, typename ScanKeyT<KM>::T &) {}
};
template <int R, int END, int KM> struct KeyStepsDiag {
	typedef typename ScanKeyT<KM>::T K;
	static S2TC_D void run(const int (&X)[16], int t, uint32_t w, uint32_t cap, K &kmin)
	{
		int s = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const int y = dpp_i<0x120 + R>(X[k]);
			s += X[k] < y ? X[k] : y;
		}
		const uint32_t b = dpp_u<0x120 + R>(w);
		K key = pair_key<KM>(s, b, cap);
		// R = 8 meets each pair twice, lanes t < 8 keep it
		if (R == 8) key = t >= 8 ? ~(K)0 : key;
		kmin = min(kmin, key);
		KeyStepsDiag<R + 1, END, KM>::run(X, t, w, cap, kmin);
	}
};
template <int END, int KM> struct KeyStepsDiag<END, END, KM> {
	static S2TC_D void run(const int (&)[16], int, uint32_t, uint32_t, typename ScanKeyT<KM>::T &) {}
};

// MAP: the rows are those of a candidate list without repeated colours,
// entry a of `ul` = {original index | colour << 16, original index of the
// colour's second occurrence (all-ones if none), row sum}; pair orders are
// formed from the original indices, and each colour that occurs twice also
// offers its pair (first, second), whose sum is its row sum (min(x, x) = x).
// Returns the winning pair order; best: the least sum (saturated at cap for
// KEY_SAT), row-uniform like the order.
template <int KM, bool MAP = false>
S2TC_D uint32_t reduce_rows_key(const int *rows, const RegRows &R, int t, int m, int pbw, const uint32_t *ul,
                                uint32_t cap, uint32_t &best)
{
	typedef typename ScanKeyT<KM>::T K;
	ScanKey st{0xffffffffu, 0xffffffffu, -1};
	const auto orig = [&](uint32_t a) -> uint32_t { return MAP ? (ul[3 * a] & 0xffffu) : a; };
	for (int p = 0; p < pbw; ++p) {
		int X[16];
		get_rows(rows, R, p, t, X);
		const int i = 16 * p + t;
		const bool irow = i < m;
		{
			K kmin = ~(K)0;
			KeyStepsDiag<1, 9, KM>::run(X, t, irow ? (uint32_t)t : 0xffffffffu, cap, kmin);
			const uint32_t b = (uint32_t)(kmin & 15u), tt = (uint32_t)t;
			const uint32_t lo = 16u * p + min(tt, b), hi = 16u * p + max(tt, b);
			const bool ok = irow && key_valid<KM>(kmin);
			merge_key((uint32_t)(kmin >> 4), ok ? (orig(lo) << 16) | orig(hi) : 0u, ok, st);
		}
		for (int q = p + 1; q < pbw; ++q) {
			int Y[16];
			get_rows(rows, R, q, t, Y);
			K kmin = ~(K)0;
			KeyStepsOff<0, 16, KM>::run(X, Y, 16 * q + t < m ? (uint32_t)t : 0xffffffffu, cap, kmin);
			const bool ok = irow && key_valid<KM>(kmin);
			const uint32_t j = 16u * q + (uint32_t)(kmin & 15u);
			merge_key((uint32_t)(kmin >> 4), ok ? (orig((uint32_t)i) << 16) | orig(j) : 0u, ok, st);
		}
		if (MAP && irow) {
			const uint32_t second = ul[3 * i + 1];
			const uint32_t rs = KM == KEY_SAT ? min(ul[3 * i + 2], cap) : ul[3 * i + 2];
			merge_key(rs, (orig((uint32_t)i) << 16) | second, second != 0xffffffffu, st);
		}
	}
	row_min_key(st.hi, st.lo);
	best = st.hi;
	return st.lo;
}

// reduce_colors_inplace / _2fixpoints argmin over the rows: returns the
// winning pair order i << 16 | j (group-uniform).  Candidate indices stay
// below 2^15 (S2TC_MAX_NRANDOM), so every order is a positive int.
template <bool NEG>
S2TC_D uint32_t reduce_rows(const int *rows, const RegRows &R, int t, int m, int pbw)
{
	ScanKey st{0xffffffffu, 0xffffffffu, -1};
	pair_scan<1, NEG>(rows, R, t, m, pbw, -1, st);
	if (NEG) {
		const int q = row_max(st.qneg);
		if (q >= 0) {
			st.hi = st.lo = 0xffffffffu;
			pair_scan<2, NEG>(rows, R, t, m, pbw, q, st);
			row_min_key(st.hi, st.lo);
			return st.lo == 0xffffffffu ? (uint32_t)q : st.lo;
		}
	}
	row_min_key(st.hi, st.lo);
	return st.lo;
}

// Per-group LDS: rows 16 S2TC_REG_RB .. mc - 1 (the first row blocks stay in
// registers), the
// features of the first 16 candidates (the columns of every row), candidates
// (565 | alpha << 16).  A row's own candidate feature is computed by its
// lane.  normal_group_smem (kernel_table.h) gives the host the same size for
// a four-block workgroup.
template <int CD> struct GroupSmem {
	uint32_t *cand;
	typename Dist<CD>::F *feat;
	int *rows;
};

template <int CD> S2TC_D size_t group_smem_bytes(int mc)
{
	return (size_t)mc * sizeof(uint32_t) + (size_t)max(mc - 16 * S2TC_REG_RB, 0) * S2TC_RS * sizeof(int) +
	       16 * sizeof(typename Dist<CD>::F);
}

template <int CD> S2TC_D GroupSmem<CD> group_smem(uint8_t *base, int g, int mc)
{
	uint8_t *p = base + (size_t)g * group_smem_bytes<CD>(mc);
	const size_t rb = (size_t)max(mc - 16 * S2TC_REG_RB, 0) * S2TC_RS * sizeof(int);
	GroupSmem<CD> s;
	s.rows = (int *)p;
	s.feat = (typename Dist<CD>::F *)(p + rb);
	s.cand = (uint32_t *)(p + rb + 16 * sizeof(typename Dist<CD>::F));
	return s;
}

// Colour refinement on 16 lanes (:767-860).  Returns this lane's 2-bit index
// (slot y*4+x), already remapped by the equal fix and the endpoint swap.
template <int CD, bool TRANS>
S2TC_D uint32_t color_refine_group(uint32_t tex, bool valid, const typename Dist<CD>::F &f, int refine,
                                   uint32_t &c0, uint32_t &c1)
{
	typedef Dist<CD> D;
	typedef typename D::F F;
	const bool transp = TRANS && texel_a(tex) == 0;
	const bool use = valid && !transp;
	if (refine == REFINE_NEVER) {
		if (TRANS ? c1 < c0 : c0 < c1) {
			uint32_t tt = c0;
			c0 = c1;
			c1 = tt;
		}
		F e0 = D::feat(c0), e1 = D::feat(c1);
		int d0 = D::d(f, e0), d1 = D::d(f, e1);
		return !valid ? 0u : transp ? 3u : (d1 < d0 ? 1u : 0u);
	}
	const uint32_t v = texel565(tex);
	const uint32_t pa = (v >> 11) | ((v & 31u) << 10) | (1u << 20);  // r | b << 10 | count << 20
	const uint32_t pg = (v >> 5) & 63u;
	const uint32_t atot = row_sum(use ? pa : 0u), gtot = row_sum(use ? pg : 0u);
	uint32_t idx = 0;
	uint32_t n0v = c0, n1v = c1, s = 0x7fffffffu;
	for (int it = 0; it < S2TC_LOOP_CAP; ++it) {
		F e0 = D::feat(n0v), e1 = D::feat(n1v);
		int d0 = D::d(f, e0), d1 = D::d(f, e1);
		const bool b = d1 < d0;
		const uint32_t score = row_sum(use ? (uint32_t)(b ? d1 : d0) : 0u);
		const uint32_t a1 = row_sum(use && b ? pa : 0u), g1 = row_sum(use && b ? pg : 0u);
		const uint32_t idx2 = !valid ? 0u : transp ? 3u : (b ? 1u : 0u);
		if (refine == REFINE_LOOP && !(score < s)) break;
		idx = idx2;
		s = score;
		c0 = n0v;
		c1 = n1v;
		const int n1 = (int)(a1 >> 20), n0 = (int)(atot >> 20) - n1;
		const uint32_t a0s = atot - a1, g0s = gtot - g1;
		bool ok = (n0 | n1) != 0;
		if (n0) {
			int r = rmean((int)(a0s & 1023u), n0) & 31, g = rmean((int)g0s, n0) & 63,
			    bb = rmean((int)((a0s >> 10) & 1023u), n0) & 31;
			n0v = (uint32_t)((r << 11) | (g << 5) | bb);
		}
		if (n1) {
			int r = rmean((int)(a1 & 1023u), n1) & 31, g = rmean((int)g1, n1) & 63,
			    bb = rmean((int)((a1 >> 10) & 1023u), n1) & 31;
			n1v = (uint32_t)((r << 11) | (g << 5) | bb);
		}
		if (refine != REFINE_LOOP) {  // ALWAYS: take the evaluated endpoints unchecked
			c0 = n0v;
			c1 = n1v;
			break;
		}
		if (!ok) break;
	}
	if (c0 == c1) {
		c1 = c0 == 0xffffu ? c1 - 1u : c1 + 1u;
		idx = idx == 1u ? 1u : 0u;
	}
	if (TRANS ? c1 < c0 : c0 < c1) {
		uint32_t tt = c0;
		c0 = c1;
		c1 = tt;
		idx ^= (idx & 2u) ? 0u : 1u;
	}
	return idx;
}

// DXT5 alpha refinement on 16 lanes (:645-765); returns this lane's 3-bit index.
S2TC_D uint32_t alpha_refine_group(int a, bool valid, int refine, int &a0, int &a1)
{
	if (refine == REFINE_NEVER) {
		if (a1 < a0) {
			int tt = a0;
			a0 = a1;
			a1 = tt;
		}
	}
	uint32_t idx = 0;
	int na0 = a0, na1 = a1;
	uint32_t s = 0x7fffffffu;
	for (int it = 0; it < S2TC_LOOP_CAP; ++it) {
		int d0 = adist(a, na0), d1 = adist(a, na1);
		bool b = d1 < d0;
		int bd = b ? d1 : d0;
		int dz = a * a, df = (255 - a) * (255 - a);
		uint32_t i2;
		int sc;
		bool acc = false;
		if (dz <= bd) {
			i2 = 6;
			sc = dz;
		} else if (df <= bd) {
			i2 = 7;
			sc = df;
		} else {
			i2 = b ? 1u : 0u;
			sc = bd;
			acc = true;
		}
		if (!valid) {
			i2 = 0;
			sc = 0;
			acc = false;
		}
		if (refine == REFINE_NEVER) return i2;
		const uint32_t score = row_sum((uint32_t)sc);
		// packed: sum a (<= 4080, 12 bits) | count << 13, per cluster
		const uint32_t c1p = row_sum(acc && b ? ((uint32_t)a | (1u << 13)) : 0u);
		const uint32_t c0p = row_sum(acc && !b ? ((uint32_t)a | (1u << 13)) : 0u);
		if (refine == REFINE_LOOP && !(score < s)) break;
		idx = i2;
		s = score;
		a0 = na0;
		a1 = na1;
		const int n0 = (int)(c0p >> 13), n1 = (int)(c1p >> 13);
		if (n0) na0 = rmean((int)(c0p & 8191u), n0) & 255;
		if (n1) na1 = rmean((int)(c1p & 8191u), n1) & 255;
		if (refine != REFINE_LOOP) {
			a0 = na0;
			a1 = na1;
			break;
		}
		if (!(n0 | n1)) break;
	}
	if (a1 == a0) {
		a1 = a0 == 255 ? a1 - 1 : a1 + 1;
		idx = idx == 1u ? 0u : idx;
	}
	if (a1 < a0) {
		int tt = a0;
		a0 = a1;
		a1 = tt;
		idx = idx == 0u ? 1u : idx == 1u ? 0u : (idx >= 6u ? idx : 7u - idx);
	}
	return idx;
}

// The whole MODE_NORMAL block on one DPP row.  tex: this lane's quantized
// texel (0 when outside the block), t: lane within the row (texel x-major
// index), pbw: row blocks = ceil((16 + max(nrandom, 0)) / 16) (kernel-uniform),
// bkey: PRNG key of the block.  Lane 0 of the row gets the block in o.
// Distances whose pair scan is a plain minimum (all but SRGB, whose wrapped
// negative sums make the fold depend on the order of every pair).
template <int CD> struct PlainScan {
	static const bool value = DistBound<CD>::value != 0 || IsSymmetric<CD>::value;
};

// DEDUP (group memory in LDS): the colour scan runs over the candidate list
// without repeated colours.  A pair's sum depends only on its two colours
// (every row is a function of its candidate's colour: d(c, c) = 0 and the
// distances are symmetric), so the first pair in the reference's (i, j)
// order with the least sum is either (first A, first B) for two colours
// A != B (every other pair of those colours comes later in that order) or
// (first A, second A), whose sum is A's row sum.  Small colour boxes (smooth
// blocks: the random candidates fall on few colours) shrink the scan from
// m (m - 1) / 2 pairs to about u^2 / 2 for u distinct colours.
template <int CD, bool DEDUP = false>
S2TC_D void encode_block_normal(uint32_t tex, int t, int w, int h, int dxt, int nrandom, int refine, uint32_t bkey,
                                int pbw, const GroupSmem<CD> &sm, BlockOut &o)
{
	typedef Dist<CD> D;
	typedef typename D::F F;
	const int lane = (int)(threadIdx.x & 63);
	const int gshift = lane & 48;
	const int x = t >> 2, y = t & 3;
	const bool valid = x < w && y < h;
	const int alpha = texel_a(tex);
	const bool use = valid && !(dxt == DXT1 && alpha == 0);
	const uint32_t v = texel565(tex);

	// gather (:940-959)
	const uint32_t mask = (uint32_t)(__ballot(use) >> gshift) & 0xffffu;
	const int pos = __builtin_popcount(mask & ((1u << t) - 1u));
	int n = __builtin_popcount(mask);
	if (use) sm.cand[pos] = v | ((uint32_t)alpha << 16);
	const bool none = n == 0;
	if (none) n = 1;
	if (none && t == 0) sm.cand[0] = 0u;
	// One colour among the candidates (the random ones are drawn inside the
	// block colours' box, :962-993, so they share it): every pair the scan can
	// pick has that colour, and the scan is skipped.  The same for DXT5 alpha,
	// except where the last-miplevel hack pairs a lone texel with alpha 0.
	const int vmin = row_min(use ? (int)v : 0xffff), vmax = row_max(use ? (int)v : 0);
	const bool cuni = none || vmin == vmax;
	bool auni = false;
	int aone = 0;
	if (dxt == DXT5) {
		const int amin = row_min(use ? alpha : 255), amax = row_max(use ? alpha : 0);
		auni = amin == amax && (nrandom > 0 || n >= 2);
		aone = amin;
	}
	__syncthreads();
	int m = n;
	int minr = 0, ming = 0, minb = 0, lr = 1, lg = 1, lb = 1;  // the colour box of the random candidates
	if (nrandom > 0) {  // (:962-993)
		int r = c_r(v), g = c_g(v), b = c_b(v);
		minr = row_min(use ? r : 255), ming = row_min(use ? g : 255), minb = row_min(use ? b : 255);
		int maxr = row_max(use ? r : 0), maxg = row_max(use ? g : 0), maxb = row_max(use ? b : 0);
		int mina = row_min(use ? alpha : 255), maxa = row_max(use ? alpha : 0);
		if (none) minr = ming = minb = maxr = maxg = maxb = mina = maxa = 0;
		lr = maxr - minr + 1, lg = maxg - ming + 1, lb = maxb - minb + 1;
		const int la = maxa - mina + 1;
		const int dpc = dxt == DXT5 ? 4 : 3;
		// a channel whose range is one value takes it for every draw (draw % 1
		// is 0): blocks of one colour (and of one alpha) skip those draws
		for (int c = t; c < nrandom; c += 16) {
			const uint32_t k = (uint32_t)(c * dpc);
			uint32_t cr = (uint32_t)minr, cg = (uint32_t)ming, cb = (uint32_t)minb, ca = 0;
			if (!cuni) {
				cr += (uint32_t)s2tc_prng_draw(bkey, k) % (uint32_t)lr;
				cg += (uint32_t)s2tc_prng_draw(bkey, k + 1) % (uint32_t)lg;
				cb += (uint32_t)s2tc_prng_draw(bkey, k + 2) % (uint32_t)lb;
			}
			if (dxt == DXT5) {
				ca = (uint32_t)mina;
				if (la > 1) ca += (uint32_t)s2tc_prng_draw(bkey, k + 3) % (uint32_t)la;
			}
			sm.cand[n + c] = (cr << 11) | (cg << 5) | cb | (ca << 16);
		}
		m = n + nrandom;
	} else if (n == 1) {  // "hack for last miplevel" (:996-1001); ca[1] pinned to 0
		if (t == 0) sm.cand[1] = sm.cand[0] & 0xffffu;
		n = m = 2;
	}
	__syncthreads();

	// Repeated colours (DEDUP).  In this block's row memory (16 (pbw -
	// S2TC_REG_RB) rows of S2TC_RS ints): T1[V] and T2[V] over the colour box
	// (V <= 256 cells) take
	// the first and second candidate index of each colour by atomic minimum,
	// then the list of first occurrences goes to the last 3 m ints, which the
	// compacted rows (written from the top) must not reach.  Every lane of the
	// workgroup passes the barriers; blocks that do not qualify skip the work.
	int mu = m, pbu = pbw;
	bool dd = false;
	uint32_t *ul = nullptr;
	if constexpr (DEDUP && PlainScan<CD>::value) {
		const int cap = 16 * (pbw - S2TC_REG_RB) * S2TC_RS;
		const int V = lr * lg * lb;
		dd = !cuni && nrandom > 0 && pbw >= 2 && V <= 256 && 2 * V + 3 * m <= cap;
		int *T1 = sm.rows, *T2 = sm.rows + V;
		ul = (uint32_t *)(sm.rows + cap - 3 * m);
		const auto cell = [&](uint32_t c) {
			return ((c_r(c) - minr) * lg + (c_g(c) - ming)) * lb + (c_b(c) - minb);
		};
		if (dd)
			for (int z = t; z < V; z += 16) T1[z] = T2[z] = 0x7fffffff;
		__syncthreads();
		if (dd)
			for (int i = t; i < m; i += 16) atomicMin(&T1[cell(sm.cand[i] & 0xffffu)], i);
		__syncthreads();
		if (dd)
			for (int i = t; i < m; i += 16) {
				const int z = cell(sm.cand[i] & 0xffffu);
				if (T1[z] != i) atomicMin(&T2[z], i);
			}
		__syncthreads();
		if (dd) {
			int base = 0;
			for (int p = 0; p < pbw; ++p) {
				const int i = 16 * p + t;
				bool first = false;
				uint32_t c = 0u;
				int sec = 0x7fffffff;
				if (i < m) {
					c = sm.cand[i] & 0xffffu;
					const int z = cell(c);
					first = T1[z] == i;
					sec = T2[z];
				}
				const uint32_t fm = (uint32_t)(__ballot(first) >> gshift) & 0xffffu;
				if (first) {
					const int a = base + __builtin_popcount(fm & ((1u << t) - 1u));
					ul[3 * a] = (uint32_t)i | (c << 16);
					ul[3 * a + 1] = sec == 0x7fffffff ? 0xffffffffu : (uint32_t)sec;
				}
				base += __builtin_popcount(fm);
			}
			mu = base;
			pbu = (mu + 15) >> 4;
			// the compacted rows (16 (pbu - S2TC_REG_RB) of them are stored)
			// must stay clear of the list, and must save work
			if (pbu >= pbw || 16 * max(pbu - S2TC_REG_RB, 0) * S2TC_RS + 3 * m > cap) {
				dd = false;
				mu = m;
				pbu = pbw;
			}
		}
		__syncthreads();
	}

	// features of the first 16 candidates (the columns)
	if (t < m) sm.feat[t] = D::feat(sm.cand[t] & 0xffffu);
	__syncthreads();

	// dists[i][k] rows (:373-392): square part dist(c[min], c[max]), box part
	// dist(c[i], c[k]); with DEDUP the rows of the list of first occurrences,
	// each with its row sum.
	RegRows R;
	for (int p = 0; p < (cuni ? 0 : pbu); ++p) {
		const int i = 16 * p + t;
		const bool irow = i < mu;
		const F fi = D::feat(dd ? (irow ? ul[3 * i] >> 16 : 0u) : sm.cand[irow ? i : 0] & 0xffffu);
		int dv[16];
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			F fk = sm.feat[k];
			int d;
			if (IsSymmetric<CD>::value) d = D::d(fi, fk);
			else d = (i < n && k < i) ? D::d(fk, fi) : D::d(fi, fk);
			dv[k] = (irow && k < n) ? d : 0;
		}
		put_rows(sm.rows, R, p, t, dv);
		if (dd && irow) {
			uint32_t rs = 0;
#pragma unroll
			for (int k = 0; k < 16; ++k) rs += (uint32_t)dv[k];
			ul[3 * i + 2] = rs;
		}
	}
	__syncthreads();
	uint32_t c0, c1;
	if (cuni) {
		c0 = c1 = none ? 0u : (uint32_t)vmin;
	} else {
		uint32_t cbest, best;
		if constexpr (DistBound<CD>::value != 0) {
			if (dd) cbest = reduce_rows_key<KEY_NARROW, true>(sm.rows, R, t, mu, pbu, ul, 0u, best);
			else cbest = reduce_rows_key<KEY_NARROW>(sm.rows, R, t, m, pbw, nullptr, 0u, best);
		} else if constexpr (IsSymmetric<CD>::value) {
			// saturated keys first; a block whose least sum reaches the cap
			// scans again with exact 64-bit keys
			const uint32_t cap = s2tc_scan_cap;
			if (dd) cbest = reduce_rows_key<KEY_SAT, true>(sm.rows, R, t, mu, pbu, ul, cap, best);
			else cbest = reduce_rows_key<KEY_SAT>(sm.rows, R, t, m, pbw, nullptr, cap, best);
			if (best >= cap) {
				if (dd) cbest = reduce_rows_key<KEY_WIDE, true>(sm.rows, R, t, mu, pbu, ul, 0u, best);
				else cbest = reduce_rows_key<KEY_WIDE>(sm.rows, R, t, m, pbw, nullptr, 0u, best);
			}
		} else {
			cbest = reduce_rows<true>(sm.rows, R, t, m, pbw);
		}
		const int ci = min((int)(cbest >> 16), m - 1), cj = min((int)(cbest & 0xffffu), m - 1);
		c0 = sm.cand[ci] & 0xffffu;
		c1 = sm.cand[cj] & 0xffffu;
	}
	int a0 = 0, a1 = 0;
	// reduce_colors_inplace_2fixpoints for DXT5 alpha (:1006, :415-478).  The
	// barriers are outside the format test: the four blocks of a workgroup
	// may have different formats in the batch kernels.
	__syncthreads();
	if (dxt == DXT5 && !auni) {
		for (int p = 0; p < pbw; ++p) {
			const int i = 16 * p + t;
			const bool irow = i < m;
			const int ai = irow ? (int)(sm.cand[i] >> 16) : 0;
			int dv[16];
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				const int ak = (int)(sm.cand[k] >> 16);
				const int f0 = adist(0, ak), f1 = adist(255, ak);
				const int fk = f1 < f0 ? f1 : f0;
				const int d = adist(ai, ak);
				dv[k] = (irow && k < n) ? (d < fk ? d : fk) : 0;
			}
			put_rows(sm.rows, R, p, t, dv);
		}
	}
	__syncthreads();
	if (dxt == DXT5 && auni) {
		a0 = a1 = aone;
	} else if (dxt == DXT5) {
		uint32_t best;
		const uint32_t abest = reduce_rows_key<KEY_NARROW>(sm.rows, R, t, m, pbw, nullptr, 0u, best);
		const int ai = min((int)(abest >> 16), m - 1), aj = min((int)(abest & 0xffffu), m - 1);
		a0 = (int)(sm.cand[ai] >> 16);
		a1 = (int)(sm.cand[aj] >> 16);
	}
	__syncthreads();

	// equal colors are BAD (:1009-1027)
	c1 = fix_equal565(c0, c1);
	if (dxt == DXT5) a1 = fix_equal_a(a0, a1);

	const F f = D::feat(v);
	const int slot = y * 4 + x;
	uint32_t cidx = (dxt == DXT1) ? color_refine_group<CD, true>(tex, valid, f, refine, c0, c1)
	                              : color_refine_group<CD, false>(tex, valid, f, refine, c0, c1);
	const uint32_t cbits = row_or(cidx << (2 * slot));
	uint32_t w0 = 0, w1 = 0;
	if (dxt == DXT3) {
		const uint64_t ab = valid ? (uint64_t)alpha << (4 * slot) : 0;
		w0 = row_or((uint32_t)ab);
		w1 = row_or((uint32_t)(ab >> 32));
	} else if (dxt == DXT5) {
		const uint32_t aidx = alpha_refine_group(alpha, valid, refine, a0, a1);
		const uint64_t ab = (uint64_t)aidx << (3 * slot);
		const uint32_t lo = row_or((uint32_t)ab), hi = row_or((uint32_t)(ab >> 32));
		const uint64_t bits = ((uint64_t)hi << 32) | lo;
		const uint64_t word = (uint64_t)(a0 & 255) | ((uint64_t)(a1 & 255) << 8) | (bits << 16);
		w0 = (uint32_t)word;
		w1 = (uint32_t)(word >> 32);
	}
	if (dxt == DXT1) {
		o.w[0] = c0 | (c1 << 16);
		o.w[1] = cbits;
		o.w[2] = o.w[3] = 0;
	} else {
		o.w[0] = w0;
		o.w[1] = w1;
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = cbits;
	}
}

}  // namespace s2tc
