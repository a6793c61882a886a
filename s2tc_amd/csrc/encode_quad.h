// encode_quad.h - quick-mode block encoder for the quadratic distances AVG,
// W0AVG and WAVG (s2tc_algorithm.cpp:217-241), one lane per block.
//
// For d(a, b) = sum_c w_c (a_c - b_c)^2 the nearest-endpoint test of
// s2tc_try_encode_block (:582-643) is linear in the texel:
//
//   d0 - d1 = 2 sum_c w_c t_c (e1_c - e0_c) - C,   C = |e1|_w^2 - |e0|_w^2
//
// so "d1 < d0" is one signed 8-bit dot product per texel (v_dot4_i32_i8 of
// the texel bytes against per-pass weight bytes, the threshold riding in its
// accumulator input).  The pass score REFINE_LOOP compares is rebuilt
// exactly from block sums (sum of |t|_w^2, the per-channel sums of all used
// texels and of those that chose endpoint 1).  The darkest/brightest search
// of MODE_FAST (:900-922) takes |t|_w^2 from v_dot4_u32_u8 and keeps
// (distance << 4 | slot) keys in one v_max/v_min each, which is the
// reference's first-occurrence rule; the two winning texels are then read
// back from the source.  Everything is exact integer arithmetic, so the
// output is bit-identical to evaluating the reference's distance texel by
// texel.
//
// Texels arrive as source words: RGBA8 (RAW) or quantized
// r5 | g6 << 8 | b5 << 16 | a << 24.
#pragma once

#include "encode_fast.h"

namespace s2tc {

template <int CD> struct QuadW { static const bool value = false; };
// weights (w_r, w_g, w_b); the texel's green byte carries AG * g so that every
// weight byte (w_c / a_c) * (e1_c - e0_c) fits in int8
template <> struct QuadW<CD_AVG> { static const bool value = true; enum { R = 4, G = 1, B = 4, AG = 1 }; };
template <> struct QuadW<CD_W0AVG> { static const bool value = true; enum { R = 1, G = 1, B = 1, AG = 1 }; };
template <> struct QuadW<CD_WAVG> { static const bool value = true; enum { R = 4, G = 4, B = 1, AG = 2 }; };

// Per-texel quantities of the quadratic encoder, from one source word.
template <int CD, bool RAW> struct QuadTexel {
	typedef QuadW<CD> W;
	uint32_t u;   // bytes (r, AG*g, b, 0)
	int d;        // |t|_w^2 = d(t, black)
	uint32_t ps;  // pack_sum: r | g << 10 | b << 21
	S2TC_D explicit QuadTexel(uint32_t s)
	{
		uint32_t rb;
		if (RAW) {
			const uint32_t l3 = s >> 3;
			rb = l3 & 0x1f001fu;
			u = rb | (W::AG == 2 ? (s >> 1) & 0x7e00u : (s >> 2) & 0x3f00u);
			ps = (l3 & 0x1fu) | (s & 0xfc00u) | ((s << 2) & 0x3e00000u);
		} else {
			rb = s & 0x1f001fu;
			u = rb | (W::AG == 2 ? (s & 0x3f00u) << 1 : s & 0x3f00u);
			ps = (s & 0x1fu) | ((s & 0x3f00u) << 2) | ((s & 0x1f0000u) << 5);
		}
		// v = the second dot operand: (w_r r, (w_g / AG) AG g, w_b b) / the texel's own factor
		uint32_t v;
		if (CD == CD_AVG) v = u + 3u * rb;           // (4r, g, 4b)
		else if (CD == CD_W0AVG) v = u;              // (r, g, b)
		else v = u + 3u * (rb & 0x1fu);              // (4r, 2g, b) against (r, 2g, b)
		d = (int)__builtin_amdgcn_udot4(u, v, 0u, false);
	}
	// DXT1: 0 for a transparent texel (quantized alpha 0), else all ones
	static S2TC_D uint32_t opaque(uint32_t s) { return RAW ? (uint32_t)((int)s >> 31) : ((s >> 24) ? ~0u : 0u); }
	static S2TC_D uint32_t to565(uint32_t s)
	{
		return RAW ? (((s >> 3) & 0x1fu) << 11) | (((s >> 10) & 0x3fu) << 5) | ((s >> 19) & 0x1fu) : texel565(s);
	}
	static S2TC_D int alpha(uint32_t s, int dxt)
	{
		const int a = (int)(s >> 24);
		return RAW ? (dxt == DXT3 ? a >> 4 : a) : a;
	}
};

S2TC_D int black565(int wr, int wg, int wb, uint32_t c)
{
	const int r = c_r(c), g = c_g(c), b = c_b(c);
	return wr * r * r + wg * g * g + wb * b * b;
}

// One s2tc_try_encode_block pass for endpoints c0, c1 (565 words).  U: the
// texels' dot bytes; ps: pack_sum of each used texel, 0 for texels that do
// not count; use2: 0b01 in the 2-bit slot of every used texel.  Returns the
// slot-ordered index bits of the used texels (1 = endpoint 1), n1 and s1 =
// count and packed sum of those texels, and the pass score sum(min(d0, d1))
// when SCORE.
template <int CD, bool SCORE>
S2TC_D uint32_t quad_pass(const uint32_t (&U)[16], const uint32_t (&ps)[16], uint32_t use2, uint32_t c0, uint32_t c1,
                          int Q, int ntot, uint32_t stot, uint32_t &score, int &n1, uint32_t &s1)
{
	typedef QuadW<CD> W;
	const int r0 = c_r(c0), g0 = c_g(c0), b0 = c_b(c0);
	const int dr = c_r(c1) - r0, dg = c_g(c1) - g0, db = c_b(c1) - b0;
	const int B0 = black565(W::R, W::G, W::B, c0);
	const int C = black565(W::R, W::G, W::B, c1) - B0;
	// negated weight bytes: sdot4(U, Wn, C >> 1) < 0  <=>  2 sum w t d > C  <=>  d1 < d0
	const uint32_t Wn = ((uint32_t)(-W::R * dr) & 0xffu) | (((uint32_t)(-(W::G / W::AG) * dg) & 0xffu) << 8) |
	                    (((uint32_t)(-W::B * db) & 0xffu) << 16);
	const int Ch = C >> 1;
	uint32_t bits = 0, s = 0;
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const int slot = (k & 3) * 4 + (k >> 2);
		const uint32_t sm = (uint32_t)(__builtin_amdgcn_sdot4((int)U[k], (int)Wn, Ch, false) >> 31);
		bits |= sm & (1u << (2 * slot));
		s += sm & ps[k];
	}
	bits &= use2;
	n1 = __builtin_popcount(bits);
	s1 = s;
	if (SCORE) {
		// sum over used texels of d0, minus sum over endpoint-1 texels of d0 - d1
		const int Sr = (int)(stot & 1023u), Sg = (int)((stot >> 10) & 2047u), Sb = (int)(stot >> 21);
		const int S1r = (int)(s & 1023u), S1g = (int)((s >> 10) & 2047u), S1b = (int)(s >> 21);
		const int d0sum = Q - 2 * (W::R * Sr * r0 + W::G * Sg * g0 + W::B * Sb * b0) + ntot * B0;
		const int gain = 2 * (W::R * dr * S1r + W::G * dg * S1g + W::B * db * S1b) - n1 * C;
		score = (uint32_t)(d0sum - gain);
	}
	return bits;
}

// colour refinement NEVER/ALWAYS/LOOP (:767-860) with the quadratic pass;
// trans3 holds index 3 in the slots of DXT1 transparent texels.
template <int CD, bool TRANS>
S2TC_D uint32_t quad_color_refine(const uint32_t (&U)[16], const uint32_t (&ps)[16], uint32_t use2, uint32_t trans3,
                                  int Q, int refine, uint32_t &c0, uint32_t &c1)
{
	const int ntot = __builtin_popcount(use2);
	uint32_t stot = 0;
#pragma unroll
	for (int k = 0; k < 16; ++k) stot += ps[k];
	uint32_t score, s1;
	int n1;
	uint32_t bits;
	if (refine == REFINE_NEVER) {
		if (TRANS ? c1 < c0 : c0 < c1) {
			uint32_t t = c0;
			c0 = c1;
			c1 = t;
		}
		return quad_pass<CD, false>(U, ps, use2, c0, c1, Q, ntot, stot, score, n1, s1) | trans3;
	}
	if (refine == REFINE_LOOP) {
		bits = 0;
		uint32_t s = 0x7fffffffu, n0v = c0, n1v = c1;
		for (int it = 0; it < S2TC_LOOP_CAP; ++it) {
			const uint32_t b2 = quad_pass<CD, true>(U, ps, use2, n0v, n1v, Q, ntot, stot, score, n1, s1);
			if (!(score < s)) break;
			bits = b2;
			s = score;
			c0 = n0v;
			c1 = n1v;
			if (!eval_color(ntot - n1, stot - s1, n1, s1, n0v, n1v)) break;
		}
	} else {
		bits = quad_pass<CD, false>(U, ps, use2, c0, c1, Q, ntot, stot, score, n1, s1);
		eval_color(ntot - n1, stot - s1, n1, s1, c0, c1);
	}
	bits |= trans3;
	color_fix_order<TRANS>(bits, c0, c1);
	return bits;
}

// The whole quick-mode block for a quadratic distance.  px[k]: source word
// of texel x = k >> 2, y = k & 3 (0 outside w x h); blk: the source at the
// block's top-left texel, rows `width` words apart.  FULL: the block is 4x4.
template <int CD, int DXT, bool FULL, bool RAW>
S2TC_D BlockOut encode_block_quad(const uint32_t (&px)[16], const uint32_t *__restrict__ blk, int width, int w, int h,
                                  int refine)
{
	typedef QuadTexel<CD, RAW> T;
	uint32_t U[16], ps[16];
	uint32_t use2 = 0, valid2 = 0;
	uint32_t kmax = 0u, kmin = 0xffffffffu;
	int Q = 0;
	int a0 = 0, a1 = 0;
	if (DXT == DXT5) a0 = a1 = T::alpha(px[0], DXT5);  // (:894-898)
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const int x = k >> 2, y = k & 3, slot = y * 4 + x;
		const bool valid = FULL || (x < w && y < h);
		const T t(px[k]);
		uint32_t usem = valid ? ~0u : 0u;
		if (DXT == DXT1) usem &= T::opaque(px[k]);
		U[k] = t.u;
		ps[k] = t.ps & usem;
		use2 |= usem & (1u << (2 * slot));
		valid2 |= valid ? 1u << (2 * slot) : 0u;
		// darkest / brightest by distance to black, first occurrence (:900-922)
		const uint32_t dx = (uint32_t)t.d & usem;
		const uint32_t dn = (usem & (uint32_t)t.d) | (~usem & 0x0fffffffu);
		Q += (int)dx;
		kmax = max(kmax, (dx << 4) | (uint32_t)(15 - k));
		kmin = min(kmin, (dn << 4) | (uint32_t)k);
		if (DXT == DXT5) {
			const int a = T::alpha(px[k], DXT5);
			const bool ta = valid && a != 255;
			a1 = (ta && a > a1) ? a : a1;
			a0 = (ta && a < a0) ? a : a0;
		}
	}
	// the winners, read back from the source (texel 0 stands in when there
	// is none; its value is then not used); no used texel leaves the dummies
	// (31,63,31) and (0,0,0) of :886-891
	const uint32_t k1 = (kmax >> 4) ? 15u - (kmax & 15u) : 0u;
	const uint32_t k0 = (kmin >> 4) != 0x0fffffffu ? (kmin & 15u) : 0u;
	const uint32_t w1 = blk[(k1 & 3u) * (uint32_t)width + (k1 >> 2)];
	const uint32_t w0 = blk[(k0 & 3u) * (uint32_t)width + (k0 >> 2)];
	uint32_t c1 = (kmax >> 4) ? T::to565(w1) : 0u;
	uint32_t c0 = (kmin >> 4) != 0x0fffffffu ? T::to565(w0) : 0xffffu;
	c1 = fix_equal565(c0, c1);
	BlockOut o;
	if (DXT == DXT1) {
		const uint32_t trans3 = (valid2 & ~use2) * 3u;
		const uint32_t bits = quad_color_refine<CD, true>(U, ps, use2, trans3, Q, refine, c0, c1);
		o.w[0] = c0 | (c1 << 16);
		o.w[1] = bits;
		o.w[2] = o.w[3] = 0;
	} else if (DXT == DXT3) {
		const uint32_t bits = quad_color_refine<CD, false>(U, ps, use2, 0u, Q, refine, c0, c1);
		uint64_t ab = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const int x = k >> 2, y = k & 3;
			const bool valid = FULL || (x < w && y < h);
			ab |= valid ? (uint64_t)T::alpha(px[k], DXT3) << (4 * (y * 4 + x)) : 0ull;
		}
		o.w[0] = (uint32_t)ab;
		o.w[1] = (uint32_t)(ab >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	} else {
		a1 = fix_equal_a(a0, a1);
		const uint32_t bits = quad_color_refine<CD, false>(U, ps, use2, 0u, Q, refine, c0, c1);
		// the alpha byte is the same in RAW and quantized DXT5 words
		uint32_t valid = 0xffffu;
		if (!FULL) {
			valid = 0;
#pragma unroll
			for (int k = 0; k < 16; ++k) valid |= ((k >> 2) < w && (k & 3) < h ? 1u : 0u) << k;
		}
		const uint64_t ab = alpha_refine_fast<FULL>(px, valid, refine, a0, a1);
		const uint64_t lo = (uint64_t)(a0 & 255) | ((uint64_t)(a1 & 255) << 8) | (ab << 16);
		o.w[0] = (uint32_t)lo;
		o.w[1] = (uint32_t)(lo >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	}
	return o;
}

template <int CD, int DXT, bool RAW>
S2TC_D BlockOut encode_block_quad_any(const uint32_t (&px)[16], const uint32_t *__restrict__ blk, int width, int w,
                                      int h, int refine)
{
	if (w == 4 && h == 4) return encode_block_quad<CD, DXT, true, RAW>(px, blk, width, 4, 4, refine);
	return encode_block_quad<CD, DXT, false, RAW>(px, blk, width, w, h, refine);
}

}  // namespace s2tc
