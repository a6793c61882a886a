/*
 * ref_harness.cpp - driver around the compiled reference encoder.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
 * with the unmodified reference sources (s2tc_algorithm.cpp,
 * s2tc_libtxc_dxtn.cpp) into oracle/_ref/libs2tc_ref.so.  It supplies:
 *
 *  - rand(): replaces libc's stream with the counter-based generator of
 *    s2tc_prng.h.  The library is linked -Bsymbolic, so the reference's
 *    calls bind here.  Two modes:
 *      per-block   the driver sets (seed, block) before each block, the draw
 *                  counter restarts at 0;
 *      per-call    inside tx_compress_dxtn the k-th call belongs to block
 *                  k / dpb (raster order), draw k % dpb, with
 *                  dpb = nrandom * (DXT5 ? 4 : 3) (s2tc_algorithm.cpp:984-992).
 *    Built with -DS2TC_STOCK_RAND (libs2tc_ref_stock.so) the override is
 *    left out and the reference draws from libc's rand().
 *  - tx_compress_dxtn: a wrapper that sets per-call mode (seed from the
 *    S2TC_RANDOM_SEED environment variable, default 0) and calls the
 *    reference's function, compiled under the name s2tc_ref_tx_compress_dxtn.
 *  - ref_encode_image_quantized[_mt]: the reference's block loop
 *    (s2tc_libtxc_dxtn.cpp:246-294) over s2tc_encode_block_func, single- or
 *    multi-threaded (dynamic block-row chunks), for golden vectors and the
 *    CPU baseline.
 */
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <strings.h>
#include <thread>
#include <vector>

#include "s2tc_algorithm.h"
#include "s2tc_prng.h"

extern "C" void s2tc_ref_tx_compress_dxtn(int srccomps, int width, int height,
                                          const unsigned char *srcPixData, unsigned int destformat,
                                          unsigned char *dest, int dstRowStride);

namespace {
struct RandState {
	int mode = 0;          // 0 per-block, 1 per-call
	uint64_t seed = 0;
	uint32_t block = 0;
	uint32_t bkey = 0;
	uint64_t counter = 0;
	uint32_t dpb = 1;
};
thread_local RandState g_rs;
std::atomic<uint64_t> g_rand_calls{0};
}

#ifdef S2TC_STOCK_RAND
/* The stock build keeps libc's rand() (libs2tc_ref_stock.so, for comparing
 * encoding error with the unmodified generator); only its seed is exposed. */
extern "C" void ref_stock_srand(unsigned int s) { srand(s); }
#else
extern "C" int rand(void)
{
	g_rand_calls.fetch_add(1, std::memory_order_relaxed);
	RandState &s = g_rs;
	if (s.mode == 1) {
		uint32_t blk = (uint32_t)(s.counter / s.dpb);
		uint32_t k = (uint32_t)(s.counter % s.dpb);
		++s.counter;
		return s2tc_prng_draw(s2tc_prng_block_key(s.seed, blk), k);
	}
	return s2tc_prng_draw(s.bkey, (uint32_t)s.counter++);
}
#endif

extern "C" uint64_t ref_rand_calls(void) { return g_rand_calls.load(); }

extern "C" void tx_compress_dxtn(int srccomps, int width, int height, const unsigned char *src,
                                 unsigned int fmt, unsigned char *dest, int dstRowStride)
{
	const char *v = getenv("S2TC_RANDOM_COLORS");
	int nrandom = v ? atoi(v) : -1;
	const char *sv = getenv("S2TC_RANDOM_SEED");
	RandState saved = g_rs;
	g_rs.mode = 1;
	g_rs.seed = sv ? strtoull(sv, nullptr, 0) : 0;
	g_rs.counter = 0;
	g_rs.dpb = (uint32_t)((nrandom > 0 ? nrandom : 1) * (fmt == 0x83F3 ? 4 : 3));
	s2tc_ref_tx_compress_dxtn(srccomps, width, height, src, fmt, dest, dstRowStride);
	g_rs = saved;
}

static void encode_rows(s2tc_encode_block_func_t f, const unsigned char *q, int width, int height,
                        int bs, int nrandom, uint64_t seed, unsigned char *dest, int row0, int row1)
{
	int bw = (width + 3) / 4;
	RandState saved = g_rs;
	g_rs.mode = 0;
	for (int by = row0; by < row1; ++by)
		for (int bx = 0; bx < bw; ++bx) {
			int i = bx * 4, j = by * 4;
			int nx = width > i + 3 ? 4 : width - i;
			int ny = height > j + 3 ? 4 : height - j;
			g_rs.bkey = s2tc_prng_block_key(seed, (uint32_t)(by * bw + bx));
			g_rs.counter = 0;
			f(dest + ((size_t)by * bw + bx) * bs, q + ((size_t)j * width + i) * 4, width, nx, ny, nrandom);
		}
	g_rs = saved;
}

extern "C" void ref_encode_image_quantized(const unsigned char *q, int width, int height, int dxt, int cd,
                                           int nrandom, int refine, uint64_t seed, unsigned char *dest,
                                           int row0, int row1)
{
	s2tc_encode_block_func_t f = s2tc_encode_block_func((DxtMode)dxt, (ColorDistMode)cd, nrandom,
	                                                    (RefinementMode)refine);
	int bs = (dxt == DXT1) ? 8 : 16;
	encode_rows(f, q, width, height, bs, nrandom, seed, dest, row0, row1);
}

/* Multi-threaded block loop; returns the thread count used. */
extern "C" int ref_encode_image_quantized_mt(const unsigned char *q, int width, int height, int dxt, int cd,
                                             int nrandom, int refine, uint64_t seed, unsigned char *dest,
                                             int nthreads)
{
	s2tc_encode_block_func_t f = s2tc_encode_block_func((DxtMode)dxt, (ColorDistMode)cd, nrandom,
	                                                    (RefinementMode)refine);
	int bs = (dxt == DXT1) ? 8 : 16;
	int bh = (height + 3) / 4;
	if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
	if (nthreads < 1) nthreads = 1;
	std::atomic<int> next{0};
	int chunk = bh / (nthreads * 16);
	if (chunk < 1) chunk = 1;
	auto work = [&]() {
		for (;;) {
			int r0 = next.fetch_add(chunk);
			if (r0 >= bh) break;
			int r1 = r0 + chunk < bh ? r0 + chunk : bh;
			encode_rows(f, q, width, height, bs, nrandom, seed, dest, r0, r1);
		}
	};
	std::vector<std::thread> th;
	for (int t = 1; t < nthreads; ++t) th.emplace_back(work);
	work();
	for (auto &t : th) t.join();
	return nthreads;
}

/* Batch of independent images (one (texture, level) each), threads over the items. */
struct RefJob {
	const unsigned char *q;
	unsigned char *dest;
	int width, height, dxt, cd, nrandom, refine;
	uint64_t seed;
};

extern "C" int ref_encode_jobs_mt(const RefJob *jobs, int njobs, int nthreads)
{
	if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
	if (nthreads < 1) nthreads = 1;
	std::atomic<int> next{0};
	auto work = [&]() {
		for (;;) {
			int i = next.fetch_add(1);
			if (i >= njobs) break;
			const RefJob &j = jobs[i];
			ref_encode_image_quantized(j.q, j.width, j.height, j.dxt, j.cd, j.nrandom, j.refine, j.seed,
			                           j.dest, 0, (j.height + 3) / 4);
		}
	};
	std::vector<std::thread> th;
	for (int t = 1; t < nthreads; ++t) th.emplace_back(work);
	work();
	for (auto &t : th) t.join();
	return nthreads;
}
