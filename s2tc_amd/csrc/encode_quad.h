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
// accumulator), whose sign is the answer.  The signs of a texel row are
// gathered into one word of 0/1 bytes with byte permutes; the cluster sums
// that s2tc_evaluate_colors_result_t needs are then dot products of that
// word with per-row byte planes of the channels (v_dot4_u32_u8), and the
// 2-bit index slots of the row come from one multiply.  The pass score
// REFINE_LOOP compares is rebuilt exactly from block sums.  The
// darkest/brightest search of MODE_FAST (:900-922) takes |t|_w^2 from
// v_dot4_u32_u8 and keeps (distance << 4 | slot) keys in one v_max/v_min
// each, which is the reference's first-occurrence rule; the two winning
// texels are read back from the source.  All of it is exact integer
// arithmetic, bit-identical to evaluating the reference's distances.
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

// Source words: RGBA8 (RAW) or quantized r5 | g6 << 8 | b5 << 16 | a << 24.
template <bool RAW> struct SrcWord {
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
	// bytes r5, g6, b5 (byte 3 is not read by the users of this word)
	static S2TC_D uint32_t quantized(uint32_t s) { return RAW ? ((s >> 3) & 0x1f001fu) | ((s >> 2) & 0x3f00u) : s; }
};

// Per-texel quantities from one source word.
template <int CD, bool RAW> struct QuadTexel : SrcWord<RAW> {
	typedef QuadW<CD> W;
	uint32_t u;  // bytes (r, AG*g, b, 0)
	int d;       // |t|_w^2 = d(t, black)
	S2TC_D explicit QuadTexel(uint32_t s)
	{
		uint32_t rb;
		if (RAW) {
			rb = (s >> 3) & 0x1f001fu;
			u = rb | (W::AG == 2 ? (s >> 1) & 0x7e00u : (s >> 2) & 0x3f00u);
		} else {
			rb = s & 0x1f001fu;
			u = rb | (W::AG == 2 ? (s & 0x3f00u) << 1 : s & 0x3f00u);
		}
		// second dot operand: bytes w_c t_c / (the texel byte's own factor)
		uint32_t v;
		if (CD == CD_AVG) v = u + 3u * rb;       // (4r, g, 4b)
		else if (CD == CD_W0AVG) v = u;          // (r, g, b)
		else v = u + 3u * (rb & 0x1fu);          // (4r, 2g, b) against (r, 2g, b)
		d = (int)__builtin_amdgcn_udot4(u, v, 0u, false);
	}
};

S2TC_D int black565(int wr, int wg, int wb, uint32_t c)
{
	const int r = c_r(c), g = c_g(c), b = c_b(c);
	return wr * r * r + wg * g * g + wb * b * b;
}

// Byte planes of the texels' dot bytes: r[y] byte x = r of texel (x, y),
// likewise AG*g and b; use[y] byte x = 1 if texel (x, y) counts.
struct QuadPlanes {
	uint32_t r[4], g[4], b[4], use[4];
};

// Texel k (x = k >> 2, y = k & 3) -> planes, with byte permutes:
// {S0, S1} selector bytes 0-3 pick S1's bytes, 4-7 S0's, 0x0c gives 0.
S2TC_D void quad_planes(const uint32_t (&U)[16], QuadPlanes &P)
{
#pragma unroll
	for (int y = 0; y < 4; ++y) {
		const uint32_t a = __builtin_amdgcn_perm(U[4 + y], U[y], 0x05010400u);       // r0 r1 g0 g1
		const uint32_t c = __builtin_amdgcn_perm(U[12 + y], U[8 + y], 0x05010400u);  // r2 r3 g2 g3
		P.r[y] = __builtin_amdgcn_perm(c, a, 0x05040100u);
		P.g[y] = __builtin_amdgcn_perm(c, a, 0x07060302u);
		P.b[y] = __builtin_amdgcn_perm(U[4 + y], U[y], 0x0c0c0602u) | __builtin_amdgcn_perm(U[12 + y], U[8 + y], 0x06020c0cu);
	}
}

// Bytes x of the result = sign bit (as 0/1) of v[x].
S2TC_D uint32_t sign_bytes(int v0, int v1, int v2, int v3)
{
	const uint32_t lo = __builtin_amdgcn_perm((uint32_t)v1, (uint32_t)v0, 0x0c0c0703u);
	const uint32_t hi = __builtin_amdgcn_perm((uint32_t)v3, (uint32_t)v2, 0x07030c0cu);
	return ((lo | hi) >> 7) & 0x01010101u;
}

// 0/1 bytes x = 0..3 -> bits 2x: one dot product with the byte weights 1, 4,
// 16, 64 (a full-rate v_dot4 where a 32-bit multiply is quarter rate)
S2TC_D uint32_t bytes_to_slots(uint32_t m) { return __builtin_amdgcn_udot4(m, 0x40100401u, 0u, false); }

struct QSums {
	int r, g, b;  // sums of the dot bytes (r, AG*g, b)
};

S2TC_D QSums plane_sums(const QuadPlanes &P, const uint32_t (&m)[4])
{
	uint32_t r = 0, g = 0, b = 0;
#pragma unroll
	for (int y = 0; y < 4; ++y) {
		r = __builtin_amdgcn_udot4(P.r[y], m[y], r, false);
		g = __builtin_amdgcn_udot4(P.g[y], m[y], g, false);
		b = __builtin_amdgcn_udot4(P.b[y], m[y], b, false);
	}
	return QSums{(int)r, (int)g, (int)b};
}

// s2tc_evaluate_colors_result_t<color_t, bigcolor_t, 1>::evaluate (:542-551)
// from sums of the dot bytes (r, AG*g, b)
template <int AG>
S2TC_D bool quad_eval_ag(int n0, const QSums &s0, int n1, const QSums &s1, uint32_t &c0, uint32_t &c1)
{
	if (!n0 && !n1) return false;
	if (n0) {
		const Recip q((uint32_t)(n0 << 1));
		const int r = (int)q.div((uint32_t)((s0.r << 1) + n0)) & 31;
		const int g = (int)q.div((uint32_t)(((s0.g / AG) << 1) + n0)) & 63;
		const int b = (int)q.div((uint32_t)((s0.b << 1) + n0)) & 31;
		c0 = (uint32_t)((r << 11) | (g << 5) | b);
	}
	if (n1) {
		const Recip q((uint32_t)(n1 << 1));
		const int r = (int)q.div((uint32_t)((s1.r << 1) + n1)) & 31;
		const int g = (int)q.div((uint32_t)(((s1.g / AG) << 1) + n1)) & 63;
		const int b = (int)q.div((uint32_t)((s1.b << 1) + n1)) & 31;
		c1 = (uint32_t)((r << 11) | (g << 5) | b);
	}
	return true;
}
template <int CD>
S2TC_D bool quad_eval(int n0, const QSums &s0, int n1, const QSums &s1, uint32_t &c0, uint32_t &c1)
{
	return quad_eval_ag<QuadW<CD>::AG>(n0, s0, n1, s1, c0, c1);
}

// One s2tc_try_encode_block pass for endpoints c0, c1 (565 words).  Returns
// the slot-ordered index bits of the used texels (1 = endpoint 1, 0 also for
// texels that do not count), n1 and s1 = count and dot-byte sums of the
// texels that chose endpoint 1, and the pass score sum(min(d0, d1)) when
// SCORE (Q = sum of |t|_w^2, tot = sums over all used texels).
template <int CD, bool SCORE>
S2TC_D uint32_t quad_pass(const uint32_t (&U)[16], const QuadPlanes &P, uint32_t c0, uint32_t c1, int Q, int ntot,
                          const QSums &tot, uint32_t &score, int &n1, QSums &s1)
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
	uint32_t bits = 0, m[4];
#pragma unroll
	for (int y = 0; y < 4; ++y) {
		const int v0 = __builtin_amdgcn_sdot4((int)U[y], (int)Wn, Ch, false);
		const int v1 = __builtin_amdgcn_sdot4((int)U[4 + y], (int)Wn, Ch, false);
		const int v2 = __builtin_amdgcn_sdot4((int)U[8 + y], (int)Wn, Ch, false);
		const int v3 = __builtin_amdgcn_sdot4((int)U[12 + y], (int)Wn, Ch, false);
		m[y] = sign_bytes(v0, v1, v2, v3) & P.use[y];
		bits |= bytes_to_slots(m[y]) << (8 * y);
	}
	n1 = __builtin_popcount(bits);
	s1 = plane_sums(P, m);
	if (SCORE) {
		// sum over used texels of d0, minus sum over endpoint-1 texels of d0 - d1
		const int d0sum = Q - 2 * (W::R * tot.r * r0 + (W::G / W::AG) * tot.g * g0 + W::B * tot.b * b0) + ntot * B0;
		const int gain = 2 * (W::R * dr * s1.r + (W::G / W::AG) * dg * s1.g + W::B * db * s1.b) - n1 * C;
		score = (uint32_t)(d0sum - gain);
	}
	return bits;
}

// colour refinement NEVER/ALWAYS/LOOP (:767-860) with the quadratic pass;
// trans3 holds index 3 in the slots of DXT1 transparent texels.
template <int CD, bool TRANS>
S2TC_D uint32_t quad_color_refine(const uint32_t (&U)[16], const QuadPlanes &P, int ntot, uint32_t trans3, int Q,
                                  int refine, uint32_t &c0, uint32_t &c1)
{
	uint32_t score;
	int n1;
	QSums s1;
	uint32_t bits;
	if (refine == REFINE_NEVER) {
		if (TRANS ? c1 < c0 : c0 < c1) {
			uint32_t t = c0;
			c0 = c1;
			c1 = t;
		}
		const QSums none{0, 0, 0};
		return quad_pass<CD, false>(U, P, c0, c1, Q, ntot, none, score, n1, s1) | trans3;
	}
	const QSums tot = plane_sums(P, P.use);
	if (refine == REFINE_LOOP) {
		bits = 0;
		uint32_t s = 0x7fffffffu, n0v = c0, n1v = c1;
		for (int it = 0; it < S2TC_LOOP_CAP; ++it) {
			const uint32_t b2 = quad_pass<CD, true>(U, P, n0v, n1v, Q, ntot, tot, score, n1, s1);
			if (!(score < s)) break;
			bits = b2;
			s = score;
			c0 = n0v;
			c1 = n1v;
			const QSums s0{tot.r - s1.r, tot.g - s1.g, tot.b - s1.b};
			if (!quad_eval<CD>(ntot - n1, s0, n1, s1, n0v, n1v)) break;
		}
	} else {
		bits = quad_pass<CD, false>(U, P, c0, c1, Q, ntot, tot, score, n1, s1);
		const QSums s0{tot.r - s1.r, tot.g - s1.g, tot.b - s1.b};
		quad_eval<CD>(ntot - n1, s0, n1, s1, c0, c1);
	}
	bits |= trans3;
	color_fix_order<TRANS>(bits, c0, c1);
	return bits;
}

// The whole quick-mode block for a quadratic distance.  px[k]: source word
// of texel x = k >> 2, y = k & 3 (0 outside w x h); blk: the source at the
// block's top-left texel, rows `width` words apart.  FULL: the block is 4x4.
// OPAQUE (DXT1): every texel is known to be opaque.
template <int CD, int DXT, bool FULL, bool RAW, bool OPAQUE>
S2TC_D BlockOut encode_block_quad(const uint32_t (&px)[16], const uint32_t *__restrict__ blk, int width, int w, int h,
                                  int refine)
{
	typedef QuadTexel<CD, RAW> T;
	constexpr bool ALL = FULL && (DXT != DXT1 || OPAQUE);  // every texel counts
	uint32_t U[16];
	QuadPlanes P;
	uint32_t usem[16];
	uint32_t kmax = 0u, kmin = 0xffffffffu;
	int Q = 0;
	int a0 = 0, a1 = 0;
	if (DXT == DXT5) a0 = a1 = T::alpha(px[0], DXT5);  // (:894-898)
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const int x = k >> 2, y = k & 3;
		const bool valid = FULL || (x < w && y < h);
		const T t(px[k]);
		uint32_t um = valid ? ~0u : 0u;
		if (DXT == DXT1 && !OPAQUE) um &= T::opaque(px[k]);
		if (ALL) um = ~0u;
		usem[k] = um;
		U[k] = t.u;
		// darkest / brightest by distance to black, first occurrence (:900-922)
		const uint32_t dx = ALL ? (uint32_t)t.d : (uint32_t)t.d & um;
		const uint32_t dn = ALL ? (uint32_t)t.d : ((um & (uint32_t)t.d) | (~um & 0x0fffffffu));
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
	quad_planes(U, P);
	uint32_t use2 = 0x55555555u, valid2 = 0x55555555u;
	if (ALL) {
#pragma unroll
		for (int y = 0; y < 4; ++y) P.use[y] = 0x01010101u;
	} else {
		use2 = 0;
#pragma unroll
		for (int y = 0; y < 4; ++y) {
			// full RAW blocks: a texel counts iff alpha >= 128, the sign bit of its word
			if (FULL && RAW)
				P.use[y] = sign_bytes((int)px[y], (int)px[4 + y], (int)px[8 + y], (int)px[12 + y]);
			else
				P.use[y] = sign_bytes((int)usem[y], (int)usem[4 + y], (int)usem[8 + y], (int)usem[12 + y]);
			use2 |= bytes_to_slots(P.use[y]) << (8 * y);
		}
		if (!FULL) {
			valid2 = 0;
#pragma unroll
			for (int k = 0; k < 16; ++k) valid2 |= ((k >> 2) < w && (k & 3) < h ? 1u : 0u) << (2 * ((k & 3) * 4 + (k >> 2)));
		}
	}
	const int ntot = __builtin_popcount(use2);
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
		const uint32_t bits = quad_color_refine<CD, true>(U, P, ntot, trans3, Q, refine, c0, c1);
		o.w[0] = c0 | (c1 << 16);
		o.w[1] = bits;
		o.w[2] = o.w[3] = 0;
	} else if (DXT == DXT3) {
		const uint32_t bits = quad_color_refine<CD, false>(U, P, ntot, 0u, Q, refine, c0, c1);
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
		const uint32_t bits = quad_color_refine<CD, false>(U, P, ntot, 0u, Q, refine, c0, c1);
		// the alpha byte is the same in RAW and quantized DXT5 words
		uint64_t lo;
		if (FULL && alpha_uniform(px)) {
			lo = alpha_uniform_half(texel_a(px[0]));
		} else {
			a1 = fix_equal_a(a0, a1);
			uint32_t valid = 0xffffu;
			if (!FULL) {
				valid = 0;
#pragma unroll
				for (int k = 0; k < 16; ++k) valid |= ((k >> 2) < w && (k & 3) < h ? 1u : 0u) << k;
			}
			const uint64_t ab = alpha_refine_fast<FULL>(px, valid, refine, a0, a1);
			lo = (uint64_t)(a0 & 255) | ((uint64_t)(a1 & 255) << 8) | (ab << 16);
		}
		o.w[0] = (uint32_t)lo;
		o.w[1] = (uint32_t)(lo >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	}
	return o;
}

// Full blocks take the 4x4 form; DXT1 waves whose blocks are all full and
// opaque (a wave-uniform vote) skip the transparency masks.
template <int CD, int DXT, bool RAW>
S2TC_D BlockOut encode_block_quad_any(const uint32_t (&px)[16], const uint32_t *__restrict__ blk, int width, int w,
                                      int h, int refine)
{
	typedef QuadTexel<CD, RAW> T;
	const bool full = w == 4 && h == 4;
	uint32_t t[16];
#pragma unroll
	for (int k = 0; k < 16; ++k) t[k] = px[k];
	if (DXT == DXT1) {
		uint32_t op = full ? ~0u : 0u;
#pragma unroll
		for (int k = 0; k < 16; ++k) op &= T::opaque(t[k]);
		if (__all(op != 0u)) {
			opaque_to_compiler(t);
			return encode_block_quad<CD, DXT, true, RAW, true>(t, blk, width, 4, 4, refine);
		}
	}
	if (full) {
		opaque_to_compiler(t);
		return encode_block_quad<CD, DXT, true, RAW, false>(t, blk, width, 4, 4, refine);
	}
	opaque_to_compiler(t);
	return encode_block_quad<CD, DXT, false, RAW, false>(t, blk, width, w, h, refine);
}

}  // namespace s2tc
