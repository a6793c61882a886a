// s2tc_device.h - device-side building blocks of the S2TC encoder (gfx950).
//
// Colours are handled as RGB565 words v = r << 11 | g << 5 | b.  The
// reference's color_t order (lexicographic r, g, b; s2tc_algorithm.cpp:70-81)
// is then the numeric order of v, its ++/-- lattice steps (:82-131) are
// v +/- 1 modulo 65536, and its packed endpoint bytes (:1046-1049) are the
// little-endian bytes of v.
//
// Distances are evaluated on per-colour "features" so that work which the
// reference repeats on every call (squares for SRGB, the sqrt luminance of
// SRGB_MIXED, the normalisation of NORMALMAP) is done once per colour; the
// remaining arithmetic is the reference's, in its order, with int32
// wrap-around where the reference overflows (SRGB).  Float code relies on
// -ffp-contract=off and on HIP's correctly rounded fp32 '/' and sqrtf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "s2tc_prng.h"

namespace s2tc {

enum { DXT1 = 0, DXT3 = 1, DXT5 = 2 };
enum { REFINE_NEVER = 0, REFINE_ALWAYS = 1, REFINE_LOOP = 2 };
enum { CD_RGB = 0, CD_YUV = 1, CD_SRGB = 2, CD_SRGB_MIXED = 3, CD_AVG = 4, CD_WAVG = 5, CD_W0AVG = 6,
       CD_NORMALMAP = 7 };

#define S2TC_D __device__ __forceinline__

S2TC_D int c_r(uint32_t v) { return (int)(v >> 11) & 31; }
S2TC_D int c_g(uint32_t v) { return (int)(v >> 5) & 63; }
S2TC_D int c_b(uint32_t v) { return (int)v & 31; }
// quantized texel (bytes r5, g6, b5, a) -> RGB565 word
S2TC_D uint32_t texel565(uint32_t t)
{
	return ((t & 0x1fu) << 11) | (((t >> 8) & 0x3fu) << 5) | ((t >> 16) & 0x1fu);
}
S2TC_D int texel_a(uint32_t t) { return (int)(t >> 24); }

// Workgroup i of a launch runs on XCD i % 8, and each of the 8 XCDs has its
// own L2.  This renumbers the G workgroups so that each XCD holds one
// contiguous run of ids (a bijection of 0..G-1): consecutive tiles, which
// share 128-byte lines of the source, then meet in one L2 instead of being
// fetched from HBM once per XCD.
S2TC_D int xcd_local_id(int bid, int G)
{
	const int q = G >> 3, r = G & 7, x = bid & 7, s = bid >> 3;
	return x * q + (x < r ? x : r) + s;
}

// SHRR(a,n) = ((a) + (1 << ((n)-1))) >> (n) with wrap-around (s2tc_algorithm.cpp:215)
S2TC_D int shrr(int a, int n) { return (int)((uint32_t)a + (1u << (n - 1))) >> n; }
S2TC_D int wmul(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }
S2TC_D int wadd(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }
S2TC_D int wsub(int a, int b) { return (int)((uint32_t)a - (uint32_t)b); }
// Full-rate 24-bit multiplies for factors known to lie in (-2^23, 2^23):
// the low 32 bits of a * b (+ c), i.e. the int32 wrap of the reference's
// expressions.  Written out because the compiler otherwise picks the
// quarter-rate 32-bit multiplies (v_mul_lo_u32, v_mad_u64_u32) for them.
// Never feed them a v_dot4 result: gfx950 needs three wait states before a
// VALU instruction reads a dot product's result, and the compiler inserts
// them only in front of instructions it generates, not in front of asm (a
// multiply one or two instructions after the dot reads the old register).
S2TC_D int mul_i24(int a, int b)
{
	int r;
	asm("v_mul_i32_i24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
	return r;
}
S2TC_D int mad_i24(int a, int b, int c)
{
	int r;
	asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
	return r;
}

struct FeatI { int x, y, z; };
struct FeatF { float x, y, z; };

template <int CD> struct Dist;

// AVG: 4dr^2 + dg^2 + 4db^2  (:217-223)
template <> struct Dist<CD_AVG> {
	typedef FeatI F;
	static const bool kCheap = true;  // featq is three bitfield extracts
	static S2TC_D F feat(uint32_t v) { return F{c_r(v), c_g(v), c_b(v)}; }
	static S2TC_D F featq(uint32_t t) { return F{(int)(t & 31u), (int)((t >> 8) & 63u), (int)((t >> 16) & 31u)}; }
	static S2TC_D int d(const F &a, const F &b)
	{
		int dr = a.x - b.x, dg = a.y - b.y, db = a.z - b.z;
		return ((dr * dr) << 2) + dg * dg + ((db * db) << 2);
	}
};
// W0AVG: dr^2 + dg^2 + db^2  (:225-232)
template <> struct Dist<CD_W0AVG> {
	typedef FeatI F;
	static const bool kCheap = true;  // featq is three bitfield extracts
	static S2TC_D F feat(uint32_t v) { return F{c_r(v), c_g(v), c_b(v)}; }
	static S2TC_D F featq(uint32_t t) { return F{(int)(t & 31u), (int)((t >> 8) & 63u), (int)((t >> 16) & 31u)}; }
	static S2TC_D int d(const F &a, const F &b)
	{
		int dr = a.x - b.x, dg = a.y - b.y, db = a.z - b.z;
		return dr * dr + dg * dg + db * db;
	}
};
// WAVG: 4dr^2 + 4dg^2 + db^2  (:234-241)
template <> struct Dist<CD_WAVG> {
	typedef FeatI F;
	static const bool kCheap = true;  // featq is three bitfield extracts
	static S2TC_D F feat(uint32_t v) { return F{c_r(v), c_g(v), c_b(v)}; }
	static S2TC_D F featq(uint32_t t) { return F{(int)(t & 31u), (int)((t >> 8) & 63u), (int)((t >> 16) & 31u)}; }
	static S2TC_D int d(const F &a, const F &b)
	{
		int dr = a.x - b.x, dg = a.y - b.y, db = a.z - b.z;
		return ((dr * dr) << 2) + ((dg * dg) << 2) + (db * db);
	}
};
// YUV / RGB: weighted Y'UV difference (:243-267).  y, u and v are linear in
// the component differences, so the feature of a colour is its own
// (Y, U, V) = (KR r + KG g + KB b, 202 r - Y, 202 b - Y) and the difference
// of two features is exactly the reference's (y, u, v).  No overflow is
// possible: |y| <= 6272 and |u|, |v| <= 202 * 31 + 6272 = 12534, so every
// term stays far below 2^31 (and every factor below 2^23).
template <int KR, int KG, int KB> struct DistYuvLike {
	typedef FeatI F;
	static const bool kCheap = true;  // a few integer ops from the texel word
	static S2TC_D F lin(int r, int g, int b)
	{
		const int y = r * KR + g * KG + b * KB;
		return F{y, r * 202 - y, b * 202 - y};
	}
	static S2TC_D F feat(uint32_t v) { return lin(c_r(v), c_g(v), c_b(v)); }
	static S2TC_D F featq(uint32_t t) { return lin((int)(t & 31u), (int)((t >> 8) & 63u), (int)((t >> 16) & 31u)); }
	static S2TC_D int d(const F &a, const F &b)
	{
		const int y = a.x - b.x, u = a.y - b.y, v = a.z - b.z;
		// (y*y << 1) + SHRR(u*u, 3) + SHRR(v*v, 4)
		return (mul_i24(y, y) << 1) + (mad_i24(u, u, 4) >> 3) + (mad_i24(v, v, 8) >> 4);
	}
};
template <> struct Dist<CD_YUV> : DistYuvLike<60, 59, 22> {};
template <> struct Dist<CD_RGB> : DistYuvLike<42, 72, 14> {};
// SRGB (:269-283): y, u and v are linear in the differences of squared
// components, so the feature of a colour is (Y, U, V) = (84 R + 72 G + 28 B,
// 409 R - Y, 409 B - Y) with R = r^2, G = g^2, B = b^2 (|Y| <= 393400, no
// overflow), and a difference of features is the reference's (y, u, v).
// SHRR(y,3)*SHRR(y,4) overflows int32 for large differences and the
// reference wraps, so the products wrap too.  Not symmetric in its
// arguments.  Every SHRR factor is below (409*961 + 393400) / 8 < 2^23.
template <> struct Dist<CD_SRGB> {
	typedef FeatI F;
	static const bool kCheap = false;  // features are kept per texel
	static S2TC_D F featq(uint32_t t) { return feat(((t & 0x1fu) << 11) | (((t >> 8) & 0x3fu) << 5) | ((t >> 16) & 0x1fu)); }
	static S2TC_D F feat(uint32_t v)
	{
		const int r = c_r(v), g = c_g(v), b = c_b(v);
		const int R = r * r, G = g * g, B = b * b;
		const int y = R * 84 + G * 72 + B * 28;
		return F{y, R * 409 - y, B * 409 - y};
	}
	static S2TC_D int d(const F &a, const F &b)
	{
		const int y = a.x - b.x, u = a.y - b.y, v = a.z - b.z;
		// SHRR(SHRR(y,3) * SHRR(y,4), 4) etc., the product and its rounding
		// term wrapping as in int32
		const int sy = mad_i24(shrr(y, 3), shrr(y, 4), 8) >> 4;
		const int su = mad_i24(shrr(u, 3), shrr(u, 4), 128) >> 8;
		const int sv = mad_i24(shrr(v, 3), shrr(v, 4), 256) >> 9;
		return wadd(wadd(sy, su), sv);
	}
};
// An endpoint e prepared for many distances d(x, e) to it (the texel
// first, as s2tc_try_encode_block calls ColorDist).  SRGB keeps e's features
// less the SHRR rounding offsets, so that SHRR(x - e, n) is one subtraction
// and one shift: (x - (e - 2^(n-1))) >> n, the subtraction wrapping like the
// reference's int arithmetic.
template <int CD> struct DistTo {
	typedef typename Dist<CD>::F F;
	F e;
	S2TC_D explicit DistTo(const F &e_) : e(e_) {}
	S2TC_D int d(const F &x) const { return Dist<CD>::d(x, e); }
};
template <> struct DistTo<CD_SRGB> {
	int y4, y8, u4, u8, v4, v8;
	S2TC_D explicit DistTo(const FeatI &e)
	    : y4(wsub(e.x, 4)), y8(wsub(e.x, 8)), u4(wsub(e.y, 4)), u8(wsub(e.y, 8)), v4(wsub(e.z, 4)), v8(wsub(e.z, 8))
	{
	}
	S2TC_D int d(const FeatI &x) const
	{
		const int sy = mad_i24(wsub(x.x, y4) >> 3, wsub(x.x, y8) >> 4, 8) >> 4;
		const int su = mad_i24(wsub(x.y, u4) >> 3, wsub(x.y, u8) >> 4, 128) >> 8;
		const int sv = mad_i24(wsub(x.z, v4) >> 3, wsub(x.z, v8) >> 4, 256) >> 9;
		return wadd(wadd(sy, su), sv);
	}
};

// srgb_get_y (:285-296): (int)(sqrtf(37*(84r^2+72g^2+28b^2)) + 0.5f); the
// argument is below 2^24 so the int->float conversion is exact.
S2TC_D int srgb_get_y(int r, int g, int b)
{
	int y = 37 * (r * r * 84 + g * g * 72 + b * b * 28);
	return (int)(sqrtf((float)y) + 0.5f);
}

#if defined(S2TC_CD) && S2TC_CD == 3
// srgb_get_y of all 65536 RGB565 colours (at most 3820), filled once per
// device by k_srgb_y_table_init; only the SRGB_MIXED kernel object defines it.
__device__ uint16_t s2tc_srgb_y_table[65536];
#define S2TC_SRGB_Y_TABLE 1
#endif

// SRGB_MIXED (:298-315): feature (Y, 191r - Y, 191b - Y); Y <= 3820, so
// the differences are below 191*31 + 2*3820 < 2^23.
template <> struct Dist<CD_SRGB_MIXED> {
	typedef FeatI F;
	static const bool kCheap = false;  // features are kept per texel
	static S2TC_D F featq(uint32_t t) { return feat(((t & 0x1fu) << 11) | (((t >> 8) & 0x3fu) << 5) | ((t >> 16) & 0x1fu)); }
	static S2TC_D int luma(uint32_t v)
	{
#ifdef S2TC_SRGB_Y_TABLE
		return (int)s2tc_srgb_y_table[v & 0xffffu];
#else
		return srgb_get_y(c_r(v), c_g(v), c_b(v));
#endif
	}
	static S2TC_D F feat(uint32_t v)
	{
		const int y = luma(v);
		return F{y, c_r(v) * 191 - y, c_b(v) * 191 - y};
	}
	static S2TC_D int d(const F &a, const F &b)
	{
		const int y = a.x - b.x, u = a.y - b.y, v = a.z - b.z;
		// (y*y << 3) + SHRR(u*u, 1) + SHRR(v*v, 2)
		return (mul_i24(y, y) << 3) + (mad_i24(u, u, 1) >> 1) + (mad_i24(v, v, 2) >> 2);
	}
};
#if defined(S2TC_CD) && S2TC_CD == 7
// NORMALMAP features of all 65536 RGB565 colours, filled once per device by
// k_nm_table_init with Dist<CD_NORMALMAP>::compute, so a lookup returns the
// very floats the formula gives.  Only the NORMALMAP kernel object defines it.
__device__ float4 s2tc_nm_table[65536];
#define S2TC_NM_TABLE 1
#endif

// NORMALMAP (:317-354): feature = the normalised vector, computed with the
// reference's operation order; distance = (int)(100000 * sum (cb-ca)^2).
template <> struct Dist<CD_NORMALMAP> {
	typedef FeatF F;
	static const bool kCheap = false;  // features are kept per texel
	static S2TC_D F featq(uint32_t t) { return feat(((t & 0x1fu) << 11) | (((t >> 8) & 0x3fu) << 5) | ((t >> 16) & 0x1fu)); }
	static S2TC_D F compute(uint32_t v)
	{
		float x = (float)c_r(v) / 31.0f * 2.0f - 1.0f;
		float y = (float)c_g(v) / 63.0f * 2.0f - 1.0f;
		float z = (float)c_b(v) / 31.0f * 2.0f - 1.0f;
		float n = x * x + y * y + z * z;
		if (n > 0) {
			n = 1.0f / sqrtf(n);
			x *= n;
			y *= n;
			z *= n;
		}
		return F{x, y, z};
	}
#ifdef S2TC_NM_TABLE
	static S2TC_D F feat(uint32_t v)
	{
		const float4 t = s2tc_nm_table[v & 0xffffu];
		return F{t.x, t.y, t.z};
	}
#else
	static S2TC_D F feat(uint32_t v) { return compute(v); }
#endif
	static S2TC_D int d(const F &a, const F &b)
	{
		float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
		return (int)(100000.0f * (dx * dx + dy * dy + dz * dz));
	}
};

// Upper bound of one distance value, for modes whose 16-term pair sums stay
// below 2^23 (0: no such bound is used).  NORMALMAP: 100000 * |b - a|^2 of
// two unit vectors (plus float slack); quadratic modes: w . (31, 63, 31)^2.
template <int CD> struct DistBound { static const uint32_t value = 0; };
template <> struct DistBound<CD_NORMALMAP> { static const uint32_t value = 400100; };
template <> struct DistBound<CD_AVG> { static const uint32_t value = 11657; };
template <> struct DistBound<CD_W0AVG> { static const uint32_t value = 5891; };
template <> struct DistBound<CD_WAVG> { static const uint32_t value = 20681; };

template <int CD> struct IsSymmetric { static const bool value = true; };

// Features of the 16 texels of a block: kept in registers for distances
// with expensive features, recomputed from the texel word otherwise (which
// saves 48 VGPRs and keeps occupancy up).
template <int CD, bool STORE = !Dist<CD>::kCheap> struct TexFeat;
template <int CD> struct TexFeat<CD, true> {
	typename Dist<CD>::F f[16];
	S2TC_D void init(const uint32_t (&tx)[16])
	{
#pragma unroll
		for (int k = 0; k < 16; ++k) f[k] = Dist<CD>::featq(tx[k]);
	}
	S2TC_D typename Dist<CD>::F get(const uint32_t (&)[16], int k) const { return f[k]; }
};
template <int CD> struct TexFeat<CD, false> {
	S2TC_D void init(const uint32_t (&)[16]) {}
	S2TC_D typename Dist<CD>::F get(const uint32_t (&tx)[16], int k) const { return Dist<CD>::featq(tx[k]); }
};
template <> struct IsSymmetric<CD_SRGB> { static const bool value = false; };

// alpha_dist (:358-361)
S2TC_D int adist(int a, int b) { return (a - b) * (a - b); }

// floor(x / d) for 0 <= x < 8192 and 1 <= d <= 32 as (x * m) >> 18 with
// m = floor(2^18 / d) + 1: m - 2^18 / d lies in (0, 1], so the product
// overshoots x / d by less than 8192 / 2^18 = 1/32 <= 1/d, and frac(x / d)
// <= 1 - 1/d, so the floor never moves.  One 24-bit multiply per quotient.
struct Recip {
	uint32_t m;
	S2TC_D explicit Recip(uint32_t d)
	{
		uint32_t q = (uint32_t)(262144.0f * __builtin_amdgcn_rcpf((float)d));
		const int r = 262144 - (int)__umul24(q, d);  // exact correction of the estimate
		q = r < 0 ? q - 1u : ((uint32_t)r >= d ? q + 1u : q);
		m = (q + 1u) & 0x3ffffu;  // <= 131073; the mask lets the product use the 24-bit multiplier
	}
	S2TC_D uint32_t div(uint32_t x) const { return ((x & 0x1fffu) * m) >> 18; }
};
// ((S << 1) + n) / (n << 1): round-half-up mean (:547-549); s <= 255 n, n <= 16
S2TC_D int rmean(int s, int n) { return (int)Recip((uint32_t)(n << 1)).div((uint32_t)((s << 1) + n)); }

// 2-bit colour index word fix-ups (:802-804, :810-812)
S2TC_D uint32_t keep_ones(uint32_t bits)
{
	uint32_t lo = bits & 0x55555555u, hi = (bits >> 1) & 0x55555555u;
	return lo & ~hi;
}
S2TC_D uint32_t swap_remap2(uint32_t bits) { return bits ^ (~(bits >> 1) & 0x55555555u); }
// 3-bit alpha index word fix-ups (:679-684, :690-704), 16 slots in 48 bits
#define S2TC_M3 0x249249249249ull
S2TC_D uint64_t keep_not_one3(uint64_t bits)
{
	uint64_t b0 = bits & S2TC_M3, b1 = (bits >> 1) & S2TC_M3, b2 = (bits >> 2) & S2TC_M3;
	uint64_t one = b0 & ~b1 & ~b2;
	return bits & ~one;
}
S2TC_D uint64_t swap_remap3(uint64_t bits)
{
	uint64_t b1 = (bits >> 1) & S2TC_M3, b2 = (bits >> 2) & S2TC_M3;
	uint64_t flip_all = b1 ^ b2;        // 2..5 -> 7 - v
	uint64_t low = ~(b1 | b2) & S2TC_M3; // 0 <-> 1
	return bits ^ (flip_all * 7u) ^ low;
}

// equal endpoints fix (:1009-1016): c1 := c1 -/+ 1 on the 565 lattice
S2TC_D uint32_t fix_equal565(uint32_t v0, uint32_t v1)
{
	if (v0 != v1) return v1;
	return v0 == 0xffffu ? v1 - 1u : ((v1 + 1u) & 0xffffu);
}
S2TC_D int fix_equal_a(int a0, int a1)
{
	if (a0 != a1) return a1;
	return a0 == 255 ? a1 - 1 : a1 + 1;
}

// True if the 16 words share their top byte (the alpha of a DXT5 texel, in
// quantized and in RGBA8 words alike).
S2TC_D bool alpha_uniform(const uint32_t (&tx)[16])
{
	uint32_t d = 0;
#pragma unroll
	for (int k = 1; k < 16; ++k) d |= tx[k] ^ tx[0];
	return (d >> 24) == 0;
}

// The DXT5 alpha half (bytes 0-7 as one word) of a full block whose 16
// alphas all equal a.  Every endpoint search yields (a, a): the quick
// minimum and maximum agree, and in the pair scans every pair sums to 0 so
// the first pair wins.  fix_equal_a (:1009-1016) makes that (a, a + 1), or
// (255, 254) for a = 255, and one pass gives every texel the same index: 0
// (the exact endpoint) for 1 <= a <= 254, the fixpoint 6 (alpha 0) for
// a = 0, the fixpoint 7 (alpha 255) for a = 255, where the order fix-up
// (:673-705) swaps the endpoints to (254, 255).  No refinement moves an
// endpoint (the texels at endpoint 0 average to a, none is at endpoint 1,
// fixpoint texels count for neither), so NEVER, ALWAYS and LOOP agree.
S2TC_D uint64_t alpha_uniform_half(int a)
{
	const uint64_t all6 = 0xDB6DB6DB6DB6ull, all7 = 0xFFFFFFFFFFFFull;
	if (a == 0) return 0x0100ull | (all6 << 16);
	if (a == 255) return 0xFFFEull | (all7 << 16);
	return (uint64_t)a | ((uint64_t)(a + 1) << 8);
}

}  // namespace s2tc
