// encode_fast.h - quick-mode ("MODE_FAST") block encoder, one lane per block.
//
// Follows s2tc_encode_block<dxt, ColorDist, MODE_FAST, refine>
// (s2tc_algorithm.cpp:872-935 and :1009-1107): darkest/brightest colour by
// distance to black, DXT5 alpha min/max over alpha != 255 texels seeded from
// texel (0,0), equal-endpoint fix, NEVER/ALWAYS/LOOP refinement, packing.
// All sixteen texels of a block live in the lane's registers; texel k is at
// x = k >> 2, y = k & 3, so unrolled loops over k visit texels x-major like
// the reference (:900-901), which fixes the first-occurrence tie-breaks.
#pragma once

#include "s2tc_device.h"

namespace s2tc {

// An upper bound on REFINE_LOOP iterations that is never reached: the loop
// only continues while the score strictly decreases, and each iteration's
// endpoints are a function of the previous iteration's 2-way partition of at
// most 16 texels, so at most 2^16 + 1 iterations can occur.
#define S2TC_LOOP_CAP 65538

struct BlockOut { uint32_t w[4]; };

// Per-block lane masks: bit k (texel x = k >> 2, y = k & 3) of `valid` is
// set if the texel is inside the block, bit k of `trans` if it is a DXT1
// transparent texel (alpha bit 0) inside the block.
struct Masks {
	uint32_t valid, trans;
};

template <int DXT>
S2TC_D Masks block_masks(const uint32_t (&tx)[16], int w, int h)
{
	Masks m{0u, 0u};
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const int x = k >> 2, y = k & 3;
		const bool v = x < w && y < h;
		m.valid |= (v ? 1u : 0u) << k;
		if (DXT == DXT1) m.trans |= ((v && texel_a(tx[k]) == 0) ? 1u : 0u) << k;
	}
	return m;
}

// packed cluster sum term of a texel: r | g << 10 | b << 21
S2TC_D uint32_t pack_sum(uint32_t t) { return (t & 0x1fu) | ((t & 0x3f00u) << 2) | ((t & 0x1f0000u) << 5); }

template <int CD, bool TRANS>
struct ColorRefineLane {
	typedef Dist<CD> D;
	typedef typename D::F F;
	// one s2tc_try_encode_block pass (:582-643) over the lane's texels,
	// without branches: texels outside the block get index 0 and do not
	// count, transparent DXT1 texels get index 3 and do not count.
	template <bool FULL>
	static S2TC_D uint32_t try_pass(const uint32_t (&tx)[16], const TexFeat<CD> &f, const Masks &mk, uint32_t c0,
	                                uint32_t c1, uint32_t &score, int &n1, uint32_t &s1)
	{
		const DistTo<CD> e0(D::feat(c0)), e1(D::feat(c1));
		uint32_t bits = 0;
		score = 0;
		n1 = 0;
		s1 = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const int x = k >> 2, y = k & 3, slot = y * 4 + x;
			const bool valid = FULL || ((mk.valid >> k) & 1u);
			const bool transp = TRANS && ((mk.trans >> k) & 1u);
			const bool use = valid && !transp;
			const F fk = f.get(tx, k);
			const int d0 = e0.d(fk), d1 = e1.d(fk);
			const bool b = d1 < d0;
			score += use ? (uint32_t)(b ? d1 : d0) : 0u;
			const bool b1 = use && b;
			n1 += b1 ? 1 : 0;
			s1 += b1 ? pack_sum(tx[k]) : 0u;
			const uint32_t idx = transp ? 3u : (b1 ? 1u : 0u);
			bits |= idx << (2 * slot);
		}
		return bits;
	}
};

// s2tc_evaluate_colors_result_t<color_t, bigcolor_t, 1>::evaluate (:542-551)
// from packed sums (r | g << 10 | b << 21).
S2TC_D bool eval_color(int n0, uint32_t s0, int n1, uint32_t s1, uint32_t &c0, uint32_t &c1)
{
	if (!n0 && !n1) return false;
	if (n0) {
		int r = rmean((int)(s0 & 1023u), n0) & 31, g = rmean((int)((s0 >> 10) & 2047u), n0) & 63,
		    b = rmean((int)(s0 >> 21), n0) & 31;
		c0 = (uint32_t)((r << 11) | (g << 5) | b);
	}
	if (n1) {
		int r = rmean((int)(s1 & 1023u), n1) & 31, g = rmean((int)((s1 >> 10) & 2047u), n1) & 63,
		    b = rmean((int)(s1 >> 21), n1) & 31;
		c1 = (uint32_t)((r << 11) | (g << 5) | b);
	}
	return true;
}

// equal fix and endpoint order of the colour refiners (:796-813)
template <bool TRANS>
S2TC_D void color_fix_order(uint32_t &bits, uint32_t &c0, uint32_t &c1)
{
	if (c0 == c1) {
		c1 = c0 == 0xffffu ? c1 - 1u : c1 + 1u;
		bits = keep_ones(bits);
	}
	if (TRANS ? c1 < c0 : c0 < c1) {
		uint32_t t = c0;
		c0 = c1;
		c1 = t;
		bits = swap_remap2(bits);
	}
}

// alpha pass of s2tc_try_encode_block with have_0_255 (:582-643), branch-free
template <bool FULL>
S2TC_D uint64_t alpha_try_lane(const uint32_t (&tx)[16], const Masks &mk, int a0, int a1, uint32_t &score, int &n0,
                               int &s0, int &n1, int &s1)
{
	uint64_t bits = 0;
	score = 0;
	n0 = s0 = n1 = s1 = 0;
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const int x = k >> 2, y = k & 3, slot = y * 4 + x;
		const bool valid = FULL || ((mk.valid >> k) & 1u);
		const int a = texel_a(tx[k]);
		const int d0 = adist(a, a0), d1 = adist(a, a1);
		const bool b = d1 < d0;
		const int bd = b ? d1 : d0;
		const int dz = a * a, df = (255 - a) * (255 - a);
		const bool z = dz <= bd, full = !z && df <= bd;
		const bool acc = valid && !z && !full;
		const uint32_t idx = z ? 6u : full ? 7u : (b ? 1u : 0u);
		score += valid ? (uint32_t)(z ? dz : full ? df : bd) : 0u;
		n1 += (acc && b) ? 1 : 0;
		s1 += (acc && b) ? a : 0;
		n0 += (acc && !b) ? 1 : 0;
		s0 += (acc && !b) ? a : 0;
		bits |= (uint64_t)(valid ? idx : 0u) << (3 * slot);
	}
	return bits;
}

S2TC_D bool eval_alpha(int n0, int s0, int n1, int s1, int &a0, int &a1)
{
	if (!n0 && !n1) return false;
	if (n0) a0 = rmean(s0, n0) & 255;
	if (n1) a1 = rmean(s1, n1) & 255;
	return true;
}

// alpha equal fix and order (:673-705)
S2TC_D void alpha_fix_order(uint64_t &bits, int &a0, int &a1)
{
	if (a1 == a0) {
		a1 = a0 == 255 ? a1 - 1 : a1 + 1;
		bits = keep_not_one3(bits);
	}
	if (a1 < a0) {
		int t = a0;
		a0 = a1;
		a1 = t;
		bits = swap_remap3(bits);
	}
}

// DXT5 alpha refinement NEVER/ALWAYS/LOOP (:645-765) on the lane's texels
template <bool FULL>
S2TC_D uint64_t alpha_refine_lane(const uint32_t (&tx)[16], const Masks &mk, int refine, int &a0, int &a1)
{
	uint32_t score;
	int n0, s0, n1, s1;
	if (refine == REFINE_NEVER) {
		if (a1 < a0) {
			int t = a0;
			a0 = a1;
			a1 = t;
		}
		return alpha_try_lane<FULL>(tx, mk, a0, a1, score, n0, s0, n1, s1);
	}
	uint64_t bits;
	if (refine == REFINE_LOOP) {
		bits = 0;
		uint32_t s = 0x7fffffffu;
		int na0 = a0, na1 = a1;
		for (int it = 0; it < S2TC_LOOP_CAP; ++it) {
			uint64_t b2 = alpha_try_lane<FULL>(tx, mk, na0, na1, score, n0, s0, n1, s1);
			if (!(score < s)) break;
			bits = b2;
			s = score;
			a0 = na0;
			a1 = na1;
			if (!eval_alpha(n0, s0, n1, s1, na0, na1)) break;
		}
	} else {
		bits = alpha_try_lane<FULL>(tx, mk, a0, a1, score, n0, s0, n1, s1);
		eval_alpha(n0, s0, n1, s1, a0, a1);
	}
	alpha_fix_order(bits, a0, a1);
	return bits;
}

// The alpha pass of alpha_try_lane without its score, in threshold form.
// For endpoints e0, e1 in [0, 255] and a texel alpha a (all integers):
//   a^2 <= (a - e)^2         <=>  e == 0   || 2a <= e
//   (255 - a)^2 <= (a - e)^2 <=>  e == 255 || 2a >= 255 + e
//   (a - e1)^2 < (a - e0)^2  <=>  (e1 - e0) (2a - e0 - e1) > 0
// so "take 0", "take 255" and "endpoint 1 is nearer" are comparisons of a
// against per-pass values.  Counts and sums travel packed as n << 16 | sum
// (sum <= 16 * 255).
template <bool FULL>
S2TC_D uint64_t alpha_try_thr(const uint32_t (&tx)[16], uint32_t valid, int a0, int a1, int &n0, int &s0, int &n1,
                              int &s1)
{
	const int zh = min(a0 == 0 ? 1023 : a0, a1 == 0 ? 1023 : a1) >> 1;                // 0 wins  <=>  a <= zh
	const int fc = (max(a0 == 255 ? 0 : 255 + a0, a1 == 255 ? 0 : 255 + a1) + 1) >> 1;  // 255 wins <=> a >= fc
	// e1 nearer <=> (e1 - e0)(2a - e0 - e1) > 0: a > floor(S / 2) when e1 > e0,
	// a < ceil(S / 2) (255 - a > 255 - ceil(S / 2)) when e1 < e0, never when equal
	const int S = a0 + a1;
	const int flip = a1 < a0 ? 255 : 0;
	const int tb = a1 > a0 ? S >> 1 : a1 < a0 ? 255 - ((S + 1) >> 1) : 255;
	uint32_t lo = 0, hi = 0, cacc = 0, c1 = 0;
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const int x = k >> 2, y = k & 3, slot = y * 4 + x;
		const bool v = FULL || ((valid >> k) & 1u);
		const int a = texel_a(tx[k]);
		const bool z = a <= zh, full = !z && a >= fc, b = (a ^ flip) > tb;
		const bool acc = v && !z && !full;
		const uint32_t idx = v ? (z ? 6u : full ? 7u : (b ? 1u : 0u)) : 0u;
		if (3 * slot + 2 < 32) {
			lo |= idx << (3 * slot);
		} else if (3 * slot >= 32) {
			hi |= idx << (3 * slot - 32);
		} else {
			lo |= idx << (3 * slot);
			hi |= idx >> (32 - 3 * slot);
		}
		const uint32_t c = (uint32_t)a | 0x10000u;
		cacc += acc ? c : 0u;
		c1 += (acc && b) ? c : 0u;
	}
	n1 = (int)(c1 >> 16);
	s1 = (int)(c1 & 0xffffu);
	n0 = (int)(cacc >> 16) - n1;
	s0 = (int)(cacc & 0xffffu) - s1;
	return ((uint64_t)hi << 32) | lo;
}

// DXT5 alpha refinement with the threshold pass for NEVER and ALWAYS (which
// need no score); LOOP keeps the scoring pass.
template <bool FULL>
S2TC_D uint64_t alpha_refine_fast(const uint32_t (&tx)[16], uint32_t valid, int refine, int &a0, int &a1)
{
	int n0, s0, n1, s1;
	if (refine == REFINE_NEVER) {
		if (a1 < a0) {
			int t = a0;
			a0 = a1;
			a1 = t;
		}
		return alpha_try_thr<FULL>(tx, valid, a0, a1, n0, s0, n1, s1);
	}
	if (refine == REFINE_LOOP) return alpha_refine_lane<FULL>(tx, Masks{valid, 0u}, refine, a0, a1);
	uint64_t bits = alpha_try_thr<FULL>(tx, valid, a0, a1, n0, s0, n1, s1);
	eval_alpha(n0, s0, n1, s1, a0, a1);
	alpha_fix_order(bits, a0, a1);
	return bits;
}

// colour refinement NEVER/ALWAYS/LOOP (:767-860) on the lane's texels
template <int CD, bool TRANS, bool FULL>
S2TC_D uint32_t color_refine_lane(const uint32_t (&tx)[16], const TexFeat<CD> &f, const Masks &mk, int refine,
                                  uint32_t &c0, uint32_t &c1)
{
	typedef ColorRefineLane<CD, TRANS> R;
	uint32_t score, s1;
	int n1;
	if (refine == REFINE_NEVER) {
		if (TRANS ? c1 < c0 : c0 < c1) {
			uint32_t t = c0;
			c0 = c1;
			c1 = t;
		}
		return R::template try_pass<FULL>(tx, f, mk, c0, c1, score, n1, s1);
	}
	// totals over the texels that enter the accumulator, fixed per block
	const uint32_t usem = (FULL ? 0xffffu : mk.valid) & ~(TRANS ? mk.trans : 0u);
	const int ntot = __builtin_popcount(usem);
	uint32_t stot = 0;
#pragma unroll
	for (int k = 0; k < 16; ++k) stot += ((usem >> k) & 1u) ? pack_sum(tx[k]) : 0u;
	uint32_t bits;
	if (refine == REFINE_LOOP) {
		bits = 0;
		uint32_t s = 0x7fffffffu, n0v = c0, n1v = c1;
		for (int it = 0; it < S2TC_LOOP_CAP; ++it) {
			uint32_t b2 = R::template try_pass<FULL>(tx, f, mk, n0v, n1v, score, n1, s1);
			if (!(score < s)) break;
			bits = b2;
			s = score;
			c0 = n0v;
			c1 = n1v;
			if (!eval_color(ntot - n1, stot - s1, n1, s1, n0v, n1v)) break;
		}
	} else {
		bits = R::template try_pass<FULL>(tx, f, mk, c0, c1, score, n1, s1);
		eval_color(ntot - n1, stot - s1, n1, s1, c0, c1);
	}
	color_fix_order<TRANS>(bits, c0, c1);
	return bits;
}

// The whole quick-mode block.  tx[k] holds the quantized texel at x = k >> 2,
// y = k & 3 (only x < w, y < h are read).  FULL: the block is 4x4.
template <int CD, int DXT, bool FULL>
S2TC_D BlockOut encode_block_fast_t(const uint32_t (&tx)[16], int w, int h, int refine)
{
	typedef Dist<CD> D;
	typedef typename D::F F;
	const Masks mk = block_masks<DXT>(tx, FULL ? 4 : w, FULL ? 4 : h);
	const F black = D::feat(0u);
	TexFeat<CD> f;
	f.init(tx);
	uint32_t c0 = 0xffffu, c1 = 0u;  // dummy (31,63,31) and (0,0,0) (:886-891)
	int dmin = 0x7fffffff, dmax = 0;
	int a0 = texel_a(tx[0]), a1 = a0;  // (:894-898)
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const bool valid = FULL || ((mk.valid >> k) & 1u);
		const bool use = valid && !(DXT == DXT1 && ((mk.trans >> k) & 1u));
		const uint32_t v = texel565(tx[k]);
		const int d = D::d(f.get(tx, k), black);
		const bool up = use && d > dmax;
		dmax = up ? d : dmax;
		c1 = up ? v : c1;
		const bool dn = use && d < dmin;
		dmin = dn ? d : dmin;
		c0 = dn ? v : c0;
		if (DXT == DXT5) {
			const int a = texel_a(tx[k]);
			const bool t = valid && a != 255;
			a1 = (t && a > a1) ? a : a1;
			a0 = (t && a < a0) ? a : a0;
		}
	}
	c1 = fix_equal565(c0, c1);
	BlockOut o;
	if (DXT == DXT1) {
		const uint32_t bits = color_refine_lane<CD, true, FULL>(tx, f, mk, refine, c0, c1);
		o.w[0] = c0 | (c1 << 16);
		o.w[1] = bits;
		o.w[2] = o.w[3] = 0;
	} else if (DXT == DXT3) {
		const uint32_t bits = color_refine_lane<CD, false, FULL>(tx, f, mk, refine, c0, c1);
		uint64_t ab = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const int x = k >> 2, y = k & 3;
			const bool valid = FULL || ((mk.valid >> k) & 1u);
			ab |= valid ? (uint64_t)texel_a(tx[k]) << (4 * (y * 4 + x)) : 0ull;
		}
		o.w[0] = (uint32_t)ab;
		o.w[1] = (uint32_t)(ab >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	} else {
		const uint32_t bits = color_refine_lane<CD, false, FULL>(tx, f, mk, refine, c0, c1);
		uint64_t lo;
		if (FULL && alpha_uniform(tx)) {
			lo = alpha_uniform_half(texel_a(tx[0]));
		} else {
			a1 = fix_equal_a(a0, a1);
			const uint64_t ab = alpha_refine_fast<FULL>(tx, mk.valid, refine, a0, a1);
			lo = (uint64_t)(a0 & 255) | ((uint64_t)(a1 & 255) << 8) | (ab << 16);
		}
		o.w[0] = (uint32_t)lo;
		o.w[1] = (uint32_t)(lo >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	}
	return o;
}

// An empty asm that claims to rewrite the words: work derived from them is
// then computed inside each branch instead of being hoisted above it, which
// would keep it live (in registers) across both paths.
S2TC_D void opaque_to_compiler(uint32_t (&tx)[16])
{
#pragma unroll
	for (int k = 0; k < 16; ++k) asm volatile("" : "+v"(tx[k]));
}

template <int CD, int DXT>
S2TC_D BlockOut encode_block_fast(const uint32_t (&tx)[16], int w, int h, int refine)
{
	uint32_t t[16];
#pragma unroll
	for (int k = 0; k < 16; ++k) t[k] = tx[k];
	if (w == 4 && h == 4) {
		opaque_to_compiler(t);
		return encode_block_fast_t<CD, DXT, true>(t, 4, 4, refine);
	}
	opaque_to_compiler(t);
	return encode_block_fast_t<CD, DXT, false>(t, w, h, refine);
}

}  // namespace s2tc
