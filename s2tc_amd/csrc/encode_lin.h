// encode_lin.h - quick-mode block encoder for RGB, YUV and SRGB_MIXED with
// REFINE_NEVER or REFINE_ALWAYS (one assignment pass), one lane per block.
//
// These distances (s2tc_algorithm.cpp:243-267, :298-315) are a quadratic
// form of a linear feature difference plus a rounding term that depends only
// on a few low bits of that difference.  With features F = (ny, u, v) of a
// colour (ny the negated luminance-like term) and s the power of two below:
//
//   s d(a, b) = Q(Fa - Fb) + corr(Fa - Fb),   Q(y, u, v) = 32 y^2 + 2 u^2 + v^2
//
//   RGB         F = (-y', u', v') = the reference's (y, u, v) / 2 of the colour,
//               s = 4, corr = 2 (u & 1) - (v & 1)
//   YUV         F = (-y, u, v) of the colour, s = 16, corr = cu[u & 3] + cv[v & 7]
//   SRGB_MIXED  F = (-Y, 191 r - Y, 191 b - Y), s = 4, corr = 2 (u & 1) - (v & 1)
//
// (checked exhaustively on the host), so the nearest-endpoint test of
// s2tc_try_encode_block (:582-643) for a texel x is the sign of
//
//   s (d1 - d0) = Q(F1) - Q(F0) + 64 ny_x (ny0 - ny1) + 4 u_x (u0 - u1)
//                 + 2 v_x (v0 - v1) + corr(Fx - F1) - corr(Fx - F0),
//
// three 24-bit multiply-adds and a per-pass correction: for the parity form
// (pu = u_x & 1, pv = v_x & 1 against the endpoints' parities) it is linear
// in (pu, pv), two more multiply-adds; for YUV it is two byte lookups in
// per-pass tables by v_perm.  |s (d1 - d0)| < 2^31, so the int32 arithmetic
// (which wraps in the products) yields it exactly.  The cluster sums, index
// slots and endpoint evaluation are those of the quadratic encoder
// (encode_quad.h); the darkest/brightest search evaluates d(x, black)
// exactly from the features and keeps (d << 4 | slot) keys (d < 2^28).
//
// REFINE_LOOP compares pass scores and stays with encode_fast.h.
#pragma once

#include "encode_quad.h"

namespace s2tc {

template <int CD> struct LinW { static const bool value = false; };
template <> struct LinW<CD_RGB> { static const bool value = true; };
template <> struct LinW<CD_YUV> { static const bool value = true; };
template <> struct LinW<CD_SRGB_MIXED> { static const bool value = true; };

struct LinF {
	int ny, u, v;
};

// The multiplies below are the 24-bit multiply-adds of s2tc_device.h (inline
// asm, so the compiler cannot fall back to the quarter-rate 32-bit
// multiplies).  None of their operands comes from a v_dot4: the compiler
// inserts the wait states a dot product's result needs before a VALU reads
// it only in front of instructions it generates itself, not in front of asm.

template <int CD> struct LinDist;

// RGB: y' = 21 r + 36 g + 7 b, u' = 101 r - y', v' = 101 b - y'.
template <> struct LinDist<CD_RGB> {
	static const bool PARITY = true;
	typedef int Extra;  // nothing kept per texel
	// q: bytes r5, g6, b5 (byte 3 ignored)
	static S2TC_D LinF feat(uint32_t q, Extra)
	{
		const int r = (int)(q & 0x1fu), g = (int)((q >> 8) & 0x3fu), b = (int)((q >> 16) & 0x1fu);
		const int ny = mad_i24(r, -21, mad_i24(g, -36, mul_i24(b, -7)));
		return LinF{ny, mad_i24(r, 101, ny), mad_i24(b, 101, ny)};
	}
	static S2TC_D Extra extra(uint32_t) { return 0; }
	static S2TC_D LinF feat565(uint32_t c)
	{
		const int r = c_r(c), g = c_g(c), b = c_b(c);
		const int y = 21 * r + 36 * g + 7 * b;
		return LinF{-y, 101 * r - y, 101 * b - y};
	}
	// d(x, black) = 8 y'^2 + SHRR(u'^2, 1) + SHRR(v'^2, 2)  (= 2 y^2 + SHRR(u^2, 3) + SHRR(v^2, 4))
	static S2TC_D int dblack(const LinF &f)
	{
		return (mul_i24(f.ny, f.ny) << 3) + (mad_i24(f.u, f.u, 1) >> 1) + (mad_i24(f.v, f.v, 2) >> 2);
	}
};

// YUV: y = 60 r + 59 g + 22 b, u = 202 r - y, v = 202 b - y.
template <> struct LinDist<CD_YUV> {
	static const bool PARITY = false;
	typedef int Extra;  // -y of the texel, kept from the search for the pass
	static S2TC_D Extra extra(uint32_t q)
	{
		const int r = (int)(q & 0x1fu), g = (int)((q >> 8) & 0x3fu), b = (int)((q >> 16) & 0x1fu);
		return mad_i24(r, -60, mad_i24(g, -59, mul_i24(b, -22)));
	}
	static S2TC_D LinF feat(uint32_t q, Extra ny)
	{
		const int r = (int)(q & 0x1fu), b = (int)((q >> 16) & 0x1fu);
		return LinF{ny, mad_i24(r, 202, ny), mad_i24(b, 202, ny)};
	}
	static S2TC_D LinF feat565(uint32_t c)
	{
		const int r = c_r(c), g = c_g(c), b = c_b(c);
		const int y = 60 * r + 59 * g + 22 * b;
		return LinF{-y, 202 * r - y, 202 * b - y};
	}
	// (y^2 << 1) + SHRR(u^2, 3) + SHRR(v^2, 4)
	static S2TC_D int dblack(const LinF &f)
	{
		return (mul_i24(f.ny, f.ny) << 1) + (mad_i24(f.u, f.u, 4) >> 3) + (mad_i24(f.v, f.v, 8) >> 4);
	}
};

// SRGB_MIXED: Y = srgb_get_y of the colour (table), u = 191 r - Y, v = 191 b - Y.
template <> struct LinDist<CD_SRGB_MIXED> {
	static const bool PARITY = true;
	typedef int Extra;  // -Y of the texel, looked up once
	static S2TC_D Extra extra(uint32_t q)
	{
		const uint32_t c = ((q & 0x1fu) << 11) | (((q >> 8) & 0x3fu) << 5) | ((q >> 16) & 0x1fu);
		return -Dist<CD_SRGB_MIXED>::luma(c);
	}
	static S2TC_D LinF feat(uint32_t q, Extra ny)
	{
		const int r = (int)(q & 0x1fu), b = (int)((q >> 16) & 0x1fu);
		return LinF{ny, mad_i24(r, 191, ny), mad_i24(b, 191, ny)};
	}
	static S2TC_D LinF feat565(uint32_t c)
	{
		const int y = Dist<CD_SRGB_MIXED>::luma(c);
		return LinF{-y, c_r(c) * 191 - y, c_b(c) * 191 - y};
	}
	// (y^2 << 3) + SHRR(u^2, 1) + SHRR(v^2, 2)
	static S2TC_D int dblack(const LinF &f)
	{
		return (mul_i24(f.ny, f.ny) << 3) + (mad_i24(f.u, f.u, 1) >> 1) + (mad_i24(f.v, f.v, 2) >> 2);
	}
};

// The extra and the endpoint features with SRGB_MIXED's luminance either from
// the table (default) or computed in registers (AR): srgb_get_y is the device
// code that fills the table, so both give the same values.  The DXT1 quick
// kernels hold 138-145 VGPRs (three waves per SIMD), too few waves to hide
// the table's dependent gathers (one per texel and endpoint): they compute
// the luminance (16384^2 quick mode 0.71 -> 0.50 ms); the DXT3 and DXT5
// forms are VALU-bound on blocks with varying alpha and keep the table.
template <int CD, bool AR>
S2TC_D typename LinDist<CD>::Extra lin_extra(uint32_t q)
{
	if constexpr (AR) return -srgb_get_y((int)(q & 0x1fu), (int)((q >> 8) & 0x3fu), (int)((q >> 16) & 0x1fu));
	else return LinDist<CD>::extra(q);
}
template <int CD, bool AR>
S2TC_D LinF lin_feat565(uint32_t c)
{
	if constexpr (AR) {
		const int y = srgb_get_y(c_r(c), c_g(c), c_b(c));
		return LinF{-y, c_r(c) * 191 - y, c_b(c) * 191 - y};
	} else {
		return LinDist<CD>::feat565(c);
	}
}

// Q(F) = 32 ny^2 + 2 u^2 + v^2 of an endpoint (below 2^31 for every distance here)
S2TC_D int lin_q(const LinF &f) { return 32 * f.ny * f.ny + 2 * f.u * f.u + f.v * f.v; }

// Byte j of the table rotated by k: T[(j - k) & (n - 1)] for a 4-byte (n = 4,
// in `lo`) or 8-byte (n = 8, {hi, lo}) table, one v_perm with per-byte
// selectors (j - k) & (n - 1).
S2TC_D uint32_t rot4(uint32_t t, int k)
{
	const uint32_t sel = (0x07060504u - (uint32_t)(k & 3) * 0x01010101u) & 0x03030303u;
	return __builtin_amdgcn_perm(t, t, sel);
}
S2TC_D void rot8(uint32_t hi, uint32_t lo, int k, uint32_t &rhi, uint32_t &rlo)
{
	const uint32_t kk = (uint32_t)(k & 7) * 0x01010101u;
	rlo = __builtin_amdgcn_perm(hi, lo, (0x0b0a0908u - kk) & 0x07070707u);
	rhi = __builtin_amdgcn_perm(hi, lo, (0x0f0e0d0cu - kk) & 0x07070707u);
}

// YUV rounding terms of 16 d, each biased by +8 so that every byte is
// positive: cu[u & 3] = {0, -2, 8, -2}, cv[v & 7] = {0, -1, -4, 7, 0, 7, -4, -1}
#define S2TC_YUV_CU 0x06100608u
#define S2TC_YUV_CV_LO 0x0f040708u
#define S2TC_YUV_CV_HI 0x07040f08u

// The per-pass part of the test for endpoints e0, e1.  With
// dny = 64 (ny0 - ny1), du = 4 (u0 - u1), dv = 2 (v0 - v1) and K = Q(F1) - Q(F0),
//   s (d1 - d0) = K + ny_x dny + u_x du + v_x dv + correction(x),
// and u_x, v_x are the texel's channels plus ny_x:
//   RGB         u = 101 r + ny, v = 101 b + ny, ny = -(21 r + 36 g + 7 b), so the
//               linear part is K + r wr + g wg + b wb in the channels;
//   YUV         u = 202 r + ny, v = 202 b + ny: K + ny_x A + r wr + b wb, A = dny + du + dv;
//   SRGB_MIXED  u = 191 r + ny, v = 191 b + ny: likewise.
// Every weight lies in (-2^23, 2^23) for every pair of endpoints (checked on
// the host), as the 24-bit multiply-adds need.
// Parity form (RGB, SRGB_MIXED): correction = (2 su - sv) + pu (-4 su) + pv (2 sv)
// with s? = (endpoint 1 parity) - (endpoint 0 parity) and pu = u_x & 1,
// pv = v_x & 1 (RGB: b & 1 and r & 1; SRGB_MIXED: (r ^ ny) & 1, (b ^ ny) & 1);
// the constant joins K.
// YUV: correction = tu[u_x & 3] + tv[v_x & 7] - 32 from per-pass byte tables
// tu = cu[(j - u1) & 3] - cu[(j - u0) & 3] + 16, likewise tv over 8 entries;
// u_x & 3 = (2 r + g + 2 b) & 3 and v_x & 7 = (4 r + 5 g + 4 b) & 7.
struct LinPass {
	int K, A, wr, wg, wb;
	int wu, wv;               // parity form
	uint32_t tu, tvlo, tvhi;  // YUV
};

template <int CD, bool AR>
S2TC_D LinPass lin_pass_setup(uint32_t c0, uint32_t c1)
{
	typedef LinDist<CD> L;
	const LinF e0 = lin_feat565<CD, AR>(c0), e1 = lin_feat565<CD, AR>(c1);
	LinPass p;
	p.K = lin_q(e1) - lin_q(e0);
	const int dny = 64 * (e0.ny - e1.ny), du = 4 * (e0.u - e1.u), dv = 2 * (e0.v - e1.v);
	p.A = dny + du + dv;
	const int k = CD == CD_RGB ? 101 : CD == CD_YUV ? 202 : 191;
	p.wr = k * du;
	p.wb = k * dv;
	p.wg = 0;
	if (CD == CD_RGB) {
		p.wr -= 21 * p.A;
		p.wg = -36 * p.A;
		p.wb -= 7 * p.A;
	}
	if constexpr (L::PARITY) {
		const int su = (e1.u & 1) - (e0.u & 1), sv = (e1.v & 1) - (e0.v & 1);
		p.K += 2 * su - sv;
		p.wu = -4 * su;
		p.wv = 2 * sv;
		p.tu = p.tvlo = p.tvhi = 0;
	} else {
		p.K -= 32;
		p.wu = p.wv = 0;
		p.tu = rot4(S2TC_YUV_CU, e1.u) + 0x10101010u - rot4(S2TC_YUV_CU, e0.u);
		uint32_t h1, l1, h0, l0;
		rot8(S2TC_YUV_CV_HI, S2TC_YUV_CV_LO, e1.v, h1, l1);
		rot8(S2TC_YUV_CV_HI, S2TC_YUV_CV_LO, e0.v, h0, l0);
		p.tvlo = l1 + 0x10101010u - l0;
		p.tvhi = h1 + 0x10101010u - h0;
	}
	return p;
}

// s (d1 - d0) for the texel word q (bytes r5, g6, b5) with its kept extra;
// the products wrap in int32 where they exceed it, the sum does not
template <int CD>
S2TC_D int lin_test(const LinPass &p, uint32_t q, typename LinDist<CD>::Extra ex)
{
	const int r = (int)(q & 0x1fu), b = (int)((q >> 16) & 0x1fu);
	int t;
	if constexpr (CD == CD_RGB) {
		const int g = (int)((q >> 8) & 0x3fu);
		t = mad_i24(r, p.wr, mad_i24(g, p.wg, mad_i24(b, p.wb, p.K)));
		t = mad_i24(b & 1, p.wu, mad_i24(r & 1, p.wv, t));
	} else if constexpr (CD == CD_SRGB_MIXED) {
		t = mad_i24(ex, p.A, mad_i24(r, p.wr, mad_i24(b, p.wb, p.K)));
		t = mad_i24((r ^ ex) & 1, p.wu, mad_i24((b ^ ex) & 1, p.wv, t));
	} else {
		t = mad_i24(ex, p.A, mad_i24(r, p.wr, mad_i24(b, p.wb, p.K)));
		const uint32_t iu = __builtin_amdgcn_udot4(q & 0x00ffffffu, 0x00020102u, 0u, false);
		const uint32_t iv = __builtin_amdgcn_udot4(q & 0x00ffffffu, 0x00040504u, 0u, false);
		const uint32_t tu = __builtin_amdgcn_perm(p.tu, p.tu, (iu & 3u) | 0x0c0c0c00u);
		const uint32_t tv = __builtin_amdgcn_perm(p.tvhi, p.tvlo, (iv & 7u) | 0x0c0c0c00u);
		t = (int)((uint32_t)t + tu + tv);
	}
	return t;
}

// One s2tc_try_encode_block pass (no score) for endpoints c0, c1: the
// slot-ordered index bits of the used texels (1 = endpoint 1), n1 and the
// byte-plane sums of the texels that chose endpoint 1.
template <int CD, bool AR>
S2TC_D uint32_t lin_pass(const uint32_t (&U)[16], const typename LinDist<CD>::Extra (&ex)[16],
                         const QuadPlanes &P, uint32_t c0, uint32_t c1, int &n1, QSums &s1)
{
	const LinPass p = lin_pass_setup<CD, AR>(c0, c1);
	uint32_t bits = 0, m[4];
#pragma unroll
	for (int y = 0; y < 4; ++y) {
		const int v0 = lin_test<CD>(p, U[y], ex[y]);
		const int v1 = lin_test<CD>(p, U[4 + y], ex[4 + y]);
		const int v2 = lin_test<CD>(p, U[8 + y], ex[8 + y]);
		const int v3 = lin_test<CD>(p, U[12 + y], ex[12 + y]);
		m[y] = sign_bytes(v0, v1, v2, v3) & P.use[y];
		bits |= bytes_to_slots(m[y]) << (8 * y);
	}
	n1 = __builtin_popcount(bits);
	s1 = plane_sums(P, m);
	return bits;
}

// colour refinement NEVER / ALWAYS (:767-860) with the linear pass; trans3
// holds index 3 in the slots of DXT1 transparent texels.
template <int CD, bool TRANS, bool AR>
S2TC_D uint32_t lin_color_refine(const uint32_t (&U)[16], const typename LinDist<CD>::Extra (&ex)[16],
                                 const QuadPlanes &P, int ntot, uint32_t trans3, int refine, uint32_t &c0,
                                 uint32_t &c1)
{
	int n1;
	QSums s1;
	if (refine == REFINE_NEVER) {
		if (TRANS ? c1 < c0 : c0 < c1) {
			uint32_t t = c0;
			c0 = c1;
			c1 = t;
		}
		return lin_pass<CD, AR>(U, ex, P, c0, c1, n1, s1) | trans3;
	}
	const QSums tot = plane_sums(P, P.use);
	uint32_t bits = lin_pass<CD, AR>(U, ex, P, c0, c1, n1, s1);
	const QSums s0{tot.r - s1.r, tot.g - s1.g, tot.b - s1.b};
	quad_eval_ag<1>(ntot - n1, s0, n1, s1, c0, c1);
	bits |= trans3;
	color_fix_order<TRANS>(bits, c0, c1);
	return bits;
}

// The whole quick-mode block (REFINE_NEVER or REFINE_ALWAYS).  px[k]: source
// word of texel x = k >> 2, y = k & 3 (0 outside w x h); blk: the source at
// the block's top-left texel, rows `width` words apart.  FULL: the block is
// 4x4.  OPAQUE (DXT1): every texel is known to be opaque.
template <int CD, int DXT, bool FULL, bool RAW, bool OPAQUE>
S2TC_D BlockOut encode_block_lin(const uint32_t (&px)[16], const uint32_t *__restrict__ blk, int width, int w, int h,
                                 int refine)
{
	typedef LinDist<CD> L;
	typedef SrcWord<RAW> S;
	constexpr bool ALL = FULL && (DXT != DXT1 || OPAQUE);  // every texel counts
	constexpr bool AR = CD == CD_SRGB_MIXED && DXT == DXT1;
	uint32_t U[16];
	typename L::Extra ex[16];
	QuadPlanes P;
	uint32_t usem[16];
	uint32_t kmax = 0u, kmin = 0xffffffffu;
	int a0 = 0, a1 = 0;
	if (DXT == DXT5) a0 = a1 = S::alpha(px[0], DXT5);  // (:894-898)
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const int x = k >> 2, y = k & 3;
		const bool valid = FULL || (x < w && y < h);
		const uint32_t q = S::quantized(px[k]);
		uint32_t um = valid ? ~0u : 0u;
		if (DXT == DXT1 && !OPAQUE) um &= S::opaque(px[k]);
		if (ALL) um = ~0u;
		usem[k] = um;
		U[k] = q;
		ex[k] = lin_extra<CD, AR>(q);
		// darkest / brightest by distance to black, first occurrence (:900-922)
		const uint32_t d = (uint32_t)L::dblack(L::feat(q, ex[k]));
		const uint32_t dx = ALL ? d : d & um;
		const uint32_t dn = ALL ? d : ((um & d) | (~um & 0x0fffffffu));
		kmax = max(kmax, (dx << 4) | (uint32_t)(15 - k));
		kmin = min(kmin, (dn << 4) | (uint32_t)k);
		if (DXT == DXT5) {
			const int a = S::alpha(px[k], DXT5);
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
	uint32_t c1 = (kmax >> 4) ? S::to565(w1) : 0u;
	uint32_t c0 = (kmin >> 4) != 0x0fffffffu ? S::to565(w0) : 0xffffu;
	c1 = fix_equal565(c0, c1);
	BlockOut o;
	if (DXT == DXT1) {
		const uint32_t trans3 = (valid2 & ~use2) * 3u;
		const uint32_t bits = lin_color_refine<CD, true, AR>(U, ex, P, ntot, trans3, refine, c0, c1);
		o.w[0] = c0 | (c1 << 16);
		o.w[1] = bits;
		o.w[2] = o.w[3] = 0;
	} else if (DXT == DXT3) {
		const uint32_t bits = lin_color_refine<CD, false, AR>(U, ex, P, ntot, 0u, refine, c0, c1);
		uint64_t ab = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const int x = k >> 2, y = k & 3;
			const bool valid = FULL || (x < w && y < h);
			ab |= valid ? (uint64_t)S::alpha(px[k], DXT3) << (4 * (y * 4 + x)) : 0ull;
		}
		o.w[0] = (uint32_t)ab;
		o.w[1] = (uint32_t)(ab >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	} else {
		const uint32_t bits = lin_color_refine<CD, false, AR>(U, ex, P, ntot, 0u, refine, c0, c1);
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
S2TC_D BlockOut encode_block_lin_any(const uint32_t (&px)[16], const uint32_t *__restrict__ blk, int width, int w,
                                     int h, int refine)
{
	typedef SrcWord<RAW> S;
	const bool full = w == 4 && h == 4;
	uint32_t t[16];
#pragma unroll
	for (int k = 0; k < 16; ++k) t[k] = px[k];
	if (DXT == DXT1) {
		uint32_t op = full ? ~0u : 0u;
#pragma unroll
		for (int k = 0; k < 16; ++k) op &= S::opaque(t[k]);
		if (__all(op != 0u)) {
			opaque_to_compiler(t);
			return encode_block_lin<CD, DXT, true, RAW, true>(t, blk, width, 4, 4, refine);
		}
	}
	if (full) {
		opaque_to_compiler(t);
		return encode_block_lin<CD, DXT, true, RAW, false>(t, blk, width, 4, 4, refine);
	}
	opaque_to_compiler(t);
	return encode_block_lin<CD, DXT, false, RAW, false>(t, blk, width, w, h, refine);
}

}  // namespace s2tc
