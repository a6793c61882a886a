// Host-only checks of library helpers, compiled with g++ against the
// library's own headers by tests/test_host_checks.py.  Prints "ok" and exits
// 0 when every check passes; otherwise prints the failing case and exits 1.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kernel_table.h"
#include "staging.h"

using s2tc::stage_slice;

// The slices of host threads 0..nt-1 cover [0, len) exactly once, in order.
static bool slices_cover(size_t len, unsigned nt)
{
	size_t next = 0;
	for (unsigned t = 0; t < nt; ++t) {
		size_t lo, hi;
		stage_slice(len, nt, t, lo, hi);
		if (hi < lo || lo > len || hi > len) return false;
		if (hi > lo) {
			if (lo != next) return false;
			if (lo % 64) return false;
			next = hi;
		}
	}
	return next == len;
}

// The Floyd-Steinberg step divides v = q * 4095 (q <= maxval) by maxval in
// {15, 31, 63} as mul_hi(v, floor(2^32 / d) + 1); exact for every v < 2^20.
static bool fs_mulhi_division_exact()
{
	for (unsigned d : {15u, 31u, 63u}) {
		const unsigned rec = 0xffffffffu / d + 1u;
		for (unsigned v = 0; v < (1u << 20); ++v)
			if ((unsigned)(((unsigned long long)v * rec) >> 32) != v / d) {
				printf("fs division: d %u v %u\n", d, v);
				return false;
			}
	}
	return true;
}

// The batch kernels split a job's block index b into (b / bw, b % bw) with
// the multiplier bw_magic(bw): q = (b * m) >> 32, then one upward correction
// when b - q bw >= bw.  Checked against the division for every b below 2^20
// at bw 1..256, and around multiples of bw (and at the top of the range) for
// every bw up to 2048 and larger ones up to the batch limit 2^23 - 1.
static bool block_row_division(unsigned b, unsigned bw)
{
	const uint32_t m = bw_magic((int)bw);
	uint32_t q = (uint32_t)(((unsigned long long)b * m) >> 32);
	uint32_t r = b - q * bw;
	if (r >= bw) {
		++q;
		r -= bw;
	}
	if (q != b / bw || r != b % bw) {
		printf("block division: b %u bw %u -> %u %u\n", b, bw, q, r);
		return false;
	}
	return true;
}

static bool block_row_division_exact()
{
	for (unsigned bw = 1; bw <= 256; ++bw)
		for (unsigned b = 0; b < (1u << 20); ++b)
			if (!block_row_division(b, bw)) return false;
	const unsigned lim = (1u << 31) - 1;
	std::vector<unsigned> bws;
	for (unsigned bw = 257; bw <= 2048; ++bw) bws.push_back(bw);
	for (unsigned bw : {4096u, 65535u, 1000003u, (1u << 23) - 1u}) bws.push_back(bw);
	for (unsigned bw : bws) {
		// rows q with q < 2^23 and q * bw + bw - 1 <= lim
		const unsigned long long qmax = std::min<unsigned long long>((1ull << 23) - 1, (lim - (bw - 1)) / bw);
		for (unsigned long long q = 0; q <= qmax; q += (q < 4096 ? 1 : 99991)) {
			const unsigned long long b0 = q * bw;
			for (unsigned long long b : {b0, b0 + 1, b0 + bw / 2, b0 + bw - 1})
				if (!block_row_division((unsigned)b, bw)) return false;
		}
		if (!block_row_division((unsigned)std::min<unsigned long long>(lim, qmax * bw + bw - 1), bw)) return false;
	}
	return true;
}

// SHRR of s2tc_algorithm.cpp:215 (no wrap in the ranges below)
static int shrr(int a, int n) { return (a + (1 << (n - 1))) >> n; }

// The quick encoder's linear endpoint test for RGB, YUV and SRGB_MIXED rests
// on these identities, checked for every component difference (RGB, YUV) or
// every term value (SRGB_MIXED) the distances can see:
//   RGB (:256-267):  4 d = 32 y'^2 + 2 u'^2 + v'^2 + 2 (u' & 1) - (v' & 1),
//     y' = 21 dr + 36 dg + 7 db, u' = 80 dr - 36 dg - 7 db,
//     v' = -21 dr - 36 dg + 94 db (the reference's y, u, v halved), with
//     u' & 1 = db & 1 and v' & 1 = dr & 1;
//   YUV (:243-254):  16 d = 32 y^2 + 2 u^2 + v^2 + cu[u & 3] + cv[v & 7];
//   SRGB_MIXED (:298-315): 4 SHRR(u^2, 1) = 2 u^2 + 2 (u & 1) and
//     4 SHRR(v^2, 2) = v^2 - (v & 1).
static const int kCu[4] = {0, -2, 8, -2};
static const int kCv[8] = {0, -1, -4, 7, 0, 7, -4, -1};

static bool linear_test_identities()
{
	for (int dr = -31; dr <= 31; ++dr)
		for (int dg = -63; dg <= 63; ++dg)
			for (int db = -31; db <= 31; ++db) {
				{
					const int y = dr * 42 + dg * 72 + db * 14, u = dr * 202 - y, v = db * 202 - y;
					const long long d = ((long long)(y * y) << 1) + shrr(u * u, 3) + shrr(v * v, 4);
					const int y2 = 21 * dr + 36 * dg + 7 * db, u2 = 80 * dr - 36 * dg - 7 * db,
					          v2 = -21 * dr - 36 * dg + 94 * db;
					const long long q = 32LL * y2 * y2 + 2LL * u2 * u2 + (long long)v2 * v2 + 2 * (u2 & 1) - (v2 & 1);
					if (4 * d != q || (u2 & 1) != (db & 1) || (v2 & 1) != (dr & 1)) {
						printf("rgb identity: %d %d %d\n", dr, dg, db);
						return false;
					}
				}
				{
					const int y = dr * 60 + dg * 59 + db * 22, u = dr * 202 - y, v = db * 202 - y;
					const long long d = ((long long)(y * y) << 1) + shrr(u * u, 3) + shrr(v * v, 4);
					const long long q = 32LL * y * y + 2LL * u * u + (long long)v * v + kCu[u & 3] + kCv[v & 7];
					if (16 * d != q) {
						printf("yuv identity: %d %d %d\n", dr, dg, db);
						return false;
					}
				}
			}
	for (int u = -20000; u <= 20000; ++u) {
		if (4LL * shrr(u * u, 1) != 2LL * u * u + 2 * (u & 1) || 4LL * shrr(u * u, 2) != (long long)u * u - (u & 1)) {
			printf("srgb_mixed identity: %d\n", u);
			return false;
		}
	}
	return true;
}

// The per-pass weights of the linear test are factors of 24-bit
// multiply-adds: every one must lie in (-2^23, 2^23) for every endpoint pair.
// They depend on the endpoints' feature differences only (dny = 64 (ny0 -
// ny1), du = 4 (u0 - u1), dv = 2 (v0 - v1), A = dny + du + dv):
//   RGB         wr = 101 du - 21 A, wg = -36 A, wb = 101 dv - 7 A
//   YUV         A, wr = 202 du, wb = 202 dv
//   SRGB_MIXED  A, wr = 191 du, wb = 191 dv (dY over every luminance difference)
static bool fits24(long long w) { return w > -(1LL << 23) && w < (1LL << 23); }

static bool linear_test_weight_ranges()
{
	for (int dr = -31; dr <= 31; ++dr)
		for (int dg = -63; dg <= 63; ++dg)
			for (int db = -31; db <= 31; ++db) {
				{
					const long long y = 21 * dr + 36 * dg + 7 * db, u = 101LL * dr - y, v = 101LL * db - y;
					const long long A = 64 * -y + 4 * u + 2 * v;
					if (!fits24(A) || !fits24(101 * 4 * u - 21 * A) || !fits24(-36 * A) || !fits24(101 * 2 * v - 7 * A)) {
						printf("rgb weights: %d %d %d\n", dr, dg, db);
						return false;
					}
				}
				{
					const long long y = 60 * dr + 59 * dg + 22 * db, u = 202LL * dr - y, v = 202LL * db - y;
					const long long A = 64 * -y + 4 * u + 2 * v;
					if (!fits24(A) || !fits24(202 * 4 * u) || !fits24(202 * 2 * v)) {
						printf("yuv weights: %d %d %d\n", dr, dg, db);
						return false;
					}
				}
			}
	// SRGB_MIXED: Y = (int)(sqrtf(37 (84 r^2 + 72 g^2 + 28 b^2)) + 0.5f) is at most that of (31, 63, 31)
	const int ymax = (int)(sqrtf((float)(37 * (84 * 961 + 72 * 3969 + 28 * 961))) + 0.5f);
	for (int dr = -31; dr <= 31; ++dr)
		for (int db = -31; db <= 31; ++db)
			for (int dy = -ymax; dy <= ymax; ++dy) {
				const long long u = 191LL * dr - dy, v = 191LL * db - dy, A = 64LL * -dy + 4 * u + 2 * v;
				if (!fits24(A) || !fits24(191 * 4 * u) || !fits24(191 * 2 * v)) {
					printf("srgb_mixed weights: %d %d %d\n", dr, db, dy);
					return false;
				}
			}
	return true;
}

int main()
{
	if (!fs_mulhi_division_exact()) return 1;
	if (!block_row_division_exact()) return 1;
	if (!linear_test_identities()) return 1;
	if (!linear_test_weight_ranges()) return 1;
	std::vector<size_t> lens;
	for (size_t n = 0; n < 2048; ++n) lens.push_back(n);
	// chunk tails of large uploads: multiples of 64 * nt plus a few bytes,
	// e.g. the last 8 MiB chunk of a 2049 x 2049 RGB source (4,206,595 bytes)
	const size_t chunk = 8u << 20;
	for (size_t k : {1u, 2u, 3u, 4u, 5u, 7u, 8u})
		for (size_t extra = 0; extra < 300; ++extra) lens.push_back(64 * k * 4096 + extra);
	lens.push_back((size_t)2049 * 2049 * 3 % chunk);
	lens.push_back((size_t)2049 * 2049 * 4 % chunk);
	lens.push_back(chunk);
	lens.push_back(chunk - 1);
	for (size_t len : lens)
		for (unsigned nt = 1; nt <= 8; ++nt)
			if (!slices_cover(len, nt)) {
				printf("stage_slice leaves a gap: len %zu, %u threads\n", len, nt);
				return 1;
			}
	printf("ok\n");
	return 0;
}
