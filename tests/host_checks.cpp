// Host-only checks of library helpers, compiled with g++ against the
// library's own headers by tests/test_host_checks.py.  Prints "ok" and exits
// 0 when every check passes; otherwise prints the failing case and exits 1.
#include <cstdio>
#include <cstdlib>
#include <vector>

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

int main()
{
	if (!fs_mulhi_division_exact()) return 1;
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
