// txc_dxtn.cpp - the reference-compatible C ABI (txc_dxtn.h, s2tc_algorithm.h).
//
// tx_compress_dxtn follows s2tc_libtxc_dxtn.cpp:142-299: environment
// configuration read on every call, GL format switch, rgb565_image, the
// raster block loop and the destination row padding rule.  The quantization
// and the block loop run on the GPU; the host uploads the source, downloads
// the blocks and lays out the destination rows.  Each call borrows a
// context (stream + device buffers) from a pool, so concurrent callers on
// different threads do not share buffers.
#include "s2tc_internal.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <strings.h>
#include <thread>
#include <vector>

#include "s2tc_algorithm.h"
#include "staging.h"
#include "txc_dxtn.h"

namespace s2tc {

struct DevBuf {
	void *p = nullptr;
	size_t cap = 0;
	int ensure(size_t n)
	{
		if (n <= cap) return 0;
		if (p) (void)hipFree(p);
		p = nullptr;
		cap = 0;
		n = (n + 4095) & ~(size_t)4095;
		if (int r = hip_check(hipMalloc(&p, n), "hipMalloc")) return r;
		cap = n;
		return 0;
	}
};

// Pinned host memory (hipHostMalloc): DMA-able staging for the host-buffer
// entry points, grown on demand and kept with the context.
struct HostBuf {
	void *p = nullptr;
	size_t cap = 0;
	int ensure(size_t n)
	{
		if (n <= cap) return 0;
		if (p) (void)hipHostFree(p);
		p = nullptr;
		cap = 0;
		n = (n + 4095) & ~(size_t)4095;
		if (int r = hip_check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc")) return r;
		cap = n;
		return 0;
	}
	uint8_t *data() const { return (uint8_t *)p; }
};

// Inputs up to this size are copied into pinned staging before the upload;
// larger ones stream through two pinned buffers of this size.
static const size_t kStageMax = 8u << 20;
// Host threads (the caller's included) that fill a staging buffer.
static const unsigned kUploadThreads = 4;

struct Ctx {
	hipStream_t stream = nullptr;
	hipStream_t copy_stream = nullptr;    // uploads of a call encoded in bands (created on first use)
	hipStream_t dither_stream = nullptr;  // its per-band quantization
	hipStream_t d2h_stream = nullptr;     // and the download of each encoded band
	DevBuf src, q, out, scratch;
	HostBuf in_stage, in_stage2, host_out;
	hipEvent_t stage_ev[2] = {nullptr, nullptr};  // last DMA out of in_stage / in_stage2
	hipEvent_t band_ev = nullptr;                 // uploaded rows handed to the dither / compute stream
	hipEvent_t dither_ev = nullptr;               // a quantized band handed to the compute stream
	hipEvent_t enc_ev = nullptr;                  // an encoded band handed to the download stream
	// per-block calls: one mapped pinned page that the kernel reads the block's
	// texels from (offset 0) and writes its encoding to (offset 64)
	uint8_t *blk_host = nullptr;
	void *blk_dev = nullptr;
};

// A large upload in chunks of kStageMax bytes through the two pinned
// buffers: while the DMA of chunk i runs, the host threads copy chunk i + 1
// into the other buffer (one host thread alone copies pageable memory well
// below the link rate).  A buffer is refilled only after the event of the
// DMA that last read it.  The DMAs go to `st`; after each one is queued,
// on_chunk (if given) learns how many leading bytes are then on their way.
static int upload_large(Ctx *c, void *dst, const void *src, size_t n, hipStream_t st,
                        const std::function<int(size_t)> &on_chunk = nullptr)
{
	int rc = c->in_stage.ensure(kStageMax);
	if (!rc) rc = c->in_stage2.ensure(kStageMax);
	for (int b = 0; b < 2 && !rc; ++b)
		if (!c->stage_ev[b])
			rc = hip_check(hipEventCreateWithFlags(&c->stage_ev[b], hipEventDisableTiming), "hipEventCreate");
	if (rc) return rc;
	const long nchunks = (long)((n + kStageMax - 1) / kStageMax);
	const unsigned hw = std::thread::hardware_concurrency();
	const unsigned nt = std::max(1u, std::min(kUploadThreads, hw ? hw : 1u));
	std::unique_ptr<std::atomic<unsigned>[]> done(new std::atomic<unsigned>[(size_t)nchunks]);
	for (long i = 0; i < nchunks; ++i) done[i].store(0);
	std::atomic<long> open{-1};  // the highest chunk the threads may copy
	std::atomic<bool> abort{false};
	uint8_t *bufs[2] = {c->in_stage.data(), c->in_stage2.data()};
	auto copy_slice = [&](long i, unsigned t) {
		const size_t off = (size_t)i * kStageMax, len = std::min(kStageMax, n - off);
		size_t lo, hi;
		stage_slice(len, nt, t, lo, hi);
		if (hi > lo) memcpy(bufs[i & 1] + lo, (const uint8_t *)src + off + lo, hi - lo);
		done[i].fetch_add(1, std::memory_order_release);
	};
	std::vector<std::thread> workers;
	for (unsigned t = 1; t < nt; ++t)
		workers.emplace_back([&, t] {
			for (long i = 0; i < nchunks; ++i) {
				while (open.load(std::memory_order_acquire) < i) {
					if (abort.load(std::memory_order_acquire)) return;
					std::this_thread::yield();
				}
				if (abort.load(std::memory_order_acquire)) return;
				copy_slice(i, t);
			}
		});
	for (long i = 0; i < nchunks && !rc; ++i) {
		rc = hip_check(hipEventSynchronize(c->stage_ev[i & 1]), "hipEventSynchronize");
		if (rc) break;
		open.store(i, std::memory_order_release);
		copy_slice(i, 0);
		while (done[i].load(std::memory_order_acquire) < nt) std::this_thread::yield();
		const size_t off = (size_t)i * kStageMax, len = std::min(kStageMax, n - off);
		rc = hip_check(hipMemcpyAsync((uint8_t *)dst + off, bufs[i & 1], len, hipMemcpyHostToDevice, st), "H2D");
		if (!rc) rc = hip_check(hipEventRecord(c->stage_ev[i & 1], st), "hipEventRecord");
		if (!rc && on_chunk) rc = on_chunk(off + len);
	}
	if (rc) abort.store(true, std::memory_order_release);
	for (auto &w : workers) w.join();
	return rc;
}

// host -> device upload of a caller's buffer on the context's stream
static int upload(Ctx *c, void *dst, const void *src, size_t n)
{
	if (n > kStageMax) return upload_large(c, dst, src, n, c->stream);
	if (int r = c->in_stage.ensure(n)) return r;
	memcpy(c->in_stage.p, src, n);
	return hip_check(hipMemcpyAsync(dst, c->in_stage.p, n, hipMemcpyHostToDevice, c->stream), "H2D");
}

static std::mutex g_pool_mu;
static std::vector<Ctx *> g_pool;

static Ctx *ctx_acquire()
{
	{
		std::lock_guard<std::mutex> lk(g_pool_mu);
		if (!g_pool.empty()) {
			Ctx *c = g_pool.back();
			g_pool.pop_back();
			return c;
		}
	}
	Ctx *c = new Ctx;
	if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
		delete c;
		return nullptr;
	}
	return c;
}

static void ctx_release(Ctx *c)
{
	std::lock_guard<std::mutex> lk(g_pool_mu);
	g_pool.push_back(c);
}

struct CtxGuard {
	Ctx *c;
	CtxGuard() : c(ctx_acquire()) {}
	~CtxGuard()
	{
		if (c) ctx_release(c);
	}
};

// Environment configuration of s2tc_libtxc_dxtn.cpp:156-216.
struct EnvConfig {
	int cd = CD_WAVG, nrandom = -1, refine = REFINE_ALWAYS, dither = DITHER_SIMPLE;
	uint64_t seed = 0;
};

static EnvConfig read_env()
{
	EnvConfig e;
	if (const char *v = getenv("S2TC_DITHER_MODE")) {
		if (!strcasecmp(v, "NONE")) e.dither = DITHER_NONE;
		else if (!strcasecmp(v, "SIMPLE")) e.dither = DITHER_SIMPLE;
		else if (!strcasecmp(v, "FLOYDSTEINBERG")) e.dither = DITHER_FLOYDSTEINBERG;
		else fprintf(stderr, "Invalid dither mode: %s\n", v);
	}
	if (const char *v = getenv("S2TC_COLORDIST_MODE")) {
		static const char *names[] = {"RGB", "YUV", "SRGB", "SRGB_MIXED", "AVG", "WAVG", "W0AVG", "NORMALMAP"};
		int found = -1;
		for (int i = 0; i < 8; ++i)
			if (!strcasecmp(v, names[i])) found = i;
		if (found >= 0) e.cd = found;
		else fprintf(stderr, "Invalid color dist mode: %s\n", v);
	}
	if (const char *v = getenv("S2TC_RANDOM_COLORS")) e.nrandom = atoi(v);
	if (const char *v = getenv("S2TC_REFINE_COLORS")) {
		if (!strcasecmp(v, "NEVER")) e.refine = REFINE_NEVER;
		else if (!strcasecmp(v, "ALWAYS")) e.refine = REFINE_ALWAYS;
		else if (!strcasecmp(v, "LOOP")) e.refine = REFINE_LOOP;
		else fprintf(stderr, "Invalid refinement mode: %s\n", v);
	}
	if (const char *v = getenv("S2TC_RANDOM_SEED")) e.seed = strtoull(v, nullptr, 0);
	return e;
}

// Rows per band of a large call (a multiple of 4: whole block rows): the
// upload of later rows overlaps the dither and encode of earlier ones.  At
// least 512 rows, and at least 2^18 blocks per band, which the one-lane-per-
// block encoders need to fill the GPU (measured: 8192-wide normal map best at
// 512 rows, 4096-wide tiles at 1024).
static int band_rows(int width)
{
	const long long r = ((1LL << 22) + width - 1) / width;
	return (int)std::max(512LL, (r + 3) & ~3LL);
}

// Large sources (more than one staging chunk, more than one band) with the
// NONE or SIMPLE dither: the uploads run on the copy stream, and each band of
// band_rows(width) rows is quantized on the dither stream and encoded on the
// compute stream as soon as its rows are uploaded (the SIMPLE scan of a band uses few
// CUs for a serial walk, so it runs beside the previous band's encode); each
// band's blocks download on a fourth stream, beside the later uploads.
// SIMPLE carries its error state from band to band on the device; the encoder
// numbers the blocks of a band from the band's first block, so the random
// candidates are those of the whole image.
static int encode_host_bands(Ctx *c, int srccomps, int width, int height, const uint8_t *src, int dxt,
                             const EncodeArgs &base, int dither, int alphabits)
{
	const int brows = band_rows(width);
	const size_t row_bytes = (size_t)width * srccomps;
	const size_t n = (size_t)width * height;
	const int bw = (width + 3) / 4, bb = block_bytes(dxt);
	const int nbands = (height + brows - 1) / brows;
	const bool fused = dither == DITHER_NONE && srccomps == 4;
	if (!fused) {
		if (int r = c->q.ensure(n * 4)) return r;
		if (int r = c->scratch.ensure(rgb565_scratch_bytes(width, brows) + 256)) return r;
	}
	if (!c->copy_stream)
		if (int r = hip_check(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking), "hipStreamCreate"))
			return r;
	if (!c->dither_stream)
		if (int r = hip_check(hipStreamCreateWithFlags(&c->dither_stream, hipStreamNonBlocking), "hipStreamCreate"))
			return r;
	if (!c->d2h_stream)
		if (int r = hip_check(hipStreamCreateWithFlags(&c->d2h_stream, hipStreamNonBlocking), "hipStreamCreate"))
			return r;
	if (!c->enc_ev)
		if (int r = hip_check(hipEventCreateWithFlags(&c->enc_ev, hipEventDisableTiming), "hipEventCreate"))
			return r;
	const size_t obytes = (size_t)bw * ((height + 3) / 4) * bb;
	if (int r = c->host_out.ensure(obytes)) return r;
	if (!c->band_ev)
		if (int r = hip_check(hipEventCreateWithFlags(&c->band_ev, hipEventDisableTiming), "hipEventCreate"))
			return r;
	if (!c->dither_ev)
		if (int r = hip_check(hipEventCreateWithFlags(&c->dither_ev, hipEventDisableTiming), "hipEventCreate"))
			return r;
	uint8_t *carry = (uint8_t *)c->scratch.p + rgb565_scratch_bytes(width, brows);
	int next = 0;
	auto on_chunk = [&](size_t uploaded) -> int {
		bool waited = false;
		while (next < nbands && (size_t)std::min(height, (next + 1) * brows) * row_bytes <= uploaded) {
			hipStream_t first = fused ? c->stream : c->dither_stream;  // the band's first user
			if (!waited) {
				if (int r = hip_check(hipEventRecord(c->band_ev, c->copy_stream), "hipEventRecord")) return r;
				if (int r = hip_check(hipStreamWaitEvent(first, c->band_ev, 0), "hipStreamWaitEvent")) return r;
				waited = true;
			}
			const int r0 = next * brows, rows = std::min(brows, height - r0);
			const uint8_t *bsrc = (const uint8_t *)c->src.p + (size_t)r0 * row_bytes;
			EncodeArgs a = base;
			a.height = rows;
			a.block_index_base = (uint32_t)((size_t)(r0 / 4) * bw);
			a.dst = (uint8_t *)c->out.p + (size_t)(r0 / 4) * bw * bb;
			if (fused) {
				a.src = bsrc;
				a.src_kind = S2TC_SRC_RGBA8;
			} else {
				uint8_t *bq = (uint8_t *)c->q.p + (size_t)r0 * width * 4;
				if (int r = launch_rgb565(bsrc, srccomps, width, rows, alphabits, dither, bq, c->scratch.p,
				                          c->dither_stream, carry, next > 0))
					return r;
				if (int r = hip_check(hipEventRecord(c->dither_ev, c->dither_stream), "hipEventRecord")) return r;
				if (int r = hip_check(hipStreamWaitEvent(c->stream, c->dither_ev, 0), "hipStreamWaitEvent")) return r;
				a.src = bq;
				a.src_kind = S2TC_SRC_QUANTIZED;
			}
			if (int r = launch_encode(a, c->stream)) return r;
			const size_t o0 = (size_t)(r0 / 4) * bw * bb, o1 = std::min(obytes, (size_t)((r0 + rows + 3) / 4) * bw * bb);
			if (int r = hip_check(hipEventRecord(c->enc_ev, c->stream), "hipEventRecord")) return r;
			if (int r = hip_check(hipStreamWaitEvent(c->d2h_stream, c->enc_ev, 0), "hipStreamWaitEvent")) return r;
			if (int r = hip_check(hipMemcpyAsync(c->host_out.data() + o0, (uint8_t *)c->out.p + o0, o1 - o0,
			                                     hipMemcpyDeviceToHost, c->d2h_stream),
			                      "D2H"))
				return r;
			++next;
		}
		return 0;
	};
	if (int r = upload_large(c, c->src.p, src, n * srccomps, c->copy_stream, on_chunk)) return r;
	if (next != nbands) return set_error(S2TC_E_ARG, "banded encode: %d of %d bands launched", next, nbands);
	return hip_check(hipStreamSynchronize(c->d2h_stream), "hipStreamSynchronize");
}

static int encode_host_image_on(Ctx *c, int srccomps, int width, int height, const uint8_t *src, int dxt, int cd,
                                int nrandom, int refine, int dither, uint64_t seed);

// Whole-image encode of host data on the context's stream; writes the
// blocks in raster order to c->host_out.  On an error every stream of the
// context is drained before the context goes back to the pool, so that no
// launch of this call still uses its buffers when the next call resizes them.
static int encode_host_image(Ctx *c, int srccomps, int width, int height, const uint8_t *src, int dxt, int cd,
                             int nrandom, int refine, int dither, uint64_t seed)
{
	const int rc = encode_host_image_on(c, srccomps, width, height, src, dxt, cd, nrandom, refine, dither, seed);
	if (rc) {
		for (hipStream_t st : {c->stream, c->copy_stream, c->dither_stream, c->d2h_stream})
			if (st) (void)hipStreamSynchronize(st);
	}
	return rc;
}

static int encode_host_image_on(Ctx *c, int srccomps, int width, int height, const uint8_t *src, int dxt, int cd,
                                int nrandom, int refine, int dither, uint64_t seed)
{
	if (srccomps != 3) srccomps = 4;
	if (dither != DITHER_NONE && dither != DITHER_FLOYDSTEINBERG) dither = DITHER_SIMPLE;
	const size_t n = (size_t)width * height;
	const size_t nb = (size_t)((width + 3) / 4) * ((height + 3) / 4);
	const size_t obytes = nb * block_bytes(dxt);
	const int alphabits = dxt == DXT1 ? 1 : dxt == DXT3 ? 4 : 8;
	if (int r = c->src.ensure(n * srccomps)) return r;
	if (int r = c->out.ensure(obytes)) return r;
	EncodeArgs a{};
	a.dst = c->out.p;
	a.width = width;
	a.height = height;
	a.dxt = dxt;
	a.cd = cd;
	a.nrandom = nrandom;
	a.refine = refine;
	a.seed = seed;
	if (dither != DITHER_FLOYDSTEINBERG && height > band_rows(width) && n * srccomps > kStageMax) {
		// the blocks are in c->host_out once the last band's download is done
		return encode_host_bands(c, srccomps, width, height, src, dxt, a, dither, alphabits);
	} else {
		if (int r = upload(c, c->src.p, src, n * srccomps)) return r;
		if (dither == DITHER_NONE && srccomps == 4) {
			a.src = c->src.p;  // quantization fused into the encoder's loads
			a.src_kind = S2TC_SRC_RGBA8;
		} else {
			if (int r = c->q.ensure(n * 4)) return r;
			if (int r = c->scratch.ensure(rgb565_scratch_bytes(width, height))) return r;
			if (int r = launch_rgb565(c->src.p, srccomps, width, height, alphabits, dither, c->q.p, c->scratch.p,
			                          c->stream))
				return r;
			a.src = c->q.p;
			a.src_kind = S2TC_SRC_QUANTIZED;
		}
		if (int r = launch_encode(a, c->stream)) return r;
	}
	if (int r = c->host_out.ensure(obytes)) return r;
	if (int r = hip_check(hipMemcpyAsync(c->host_out.data(), c->out.p, obytes, hipMemcpyDeviceToHost, c->stream),
	                      "D2H"))
		return r;
	return hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

// ---- per-block entry points of s2tc_encode_block_func ----

static thread_local uint32_t t_block_index = 0;

static void encode_one_block(unsigned char *out, const unsigned char *rgba, int iw, int w, int h, int nrandom,
                             int dxt, int cd, int fast, int refine)
{
	CtxGuard g;
	Ctx *c = g.c;
	const uint32_t bidx = t_block_index++;
	if (!c) {
		fprintf(stderr, "s2tc: cannot create a HIP stream\n");
		return;
	}
	if (w < 1 || h < 1 || w > 4 || h > 4) return;
	const int bs = block_bytes(dxt);
	int r = 0;
	if (!c->blk_host) {
		r = hip_check(hipHostMalloc((void **)&c->blk_host, 4096, hipHostMallocMapped), "hipHostMalloc");
		if (!r) r = hip_check(hipHostGetDevicePointer(&c->blk_dev, c->blk_host, 0), "hipHostGetDevicePointer");
		if (r && c->blk_host) {
			(void)hipHostFree(c->blk_host);
			c->blk_host = nullptr;
		}
	}
	if (!r) {
		// the block's rows, packed, where the kernel reads them over the bus
		for (int y = 0; y < h; ++y) memcpy(c->blk_host + y * w * 4, rgba + (size_t)y * iw * 4, (size_t)w * 4);
		// A one-block image whose only block has index bidx in the generator.
		// The selection-time mode is kept: quick mode runs with nrandom < 0,
		// MODE_NORMAL with max(nrandom, 0) extra candidates.
		uint8_t *dev = (uint8_t *)c->blk_dev;
		EncodeArgs a{dev, dev + 64, S2TC_SRC_QUANTIZED, w, h, dxt, cd, fast ? -1 : (nrandom > 0 ? nrandom : 0),
		             refine, 0, 0};
		const char *sv = getenv("S2TC_RANDOM_SEED");
		a.seed = sv ? strtoull(sv, nullptr, 0) : 0;
		a.block_index_base = bidx;
		r = launch_encode(a, c->stream);
	}
	if (!r) r = hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
	if (r) {
		fprintf(stderr, "s2tc: %s\n", s2tc_amd_last_error());
		return;
	}
	memcpy(out, c->blk_host + 64, bs);
}

template <int DXT, int CD, int FAST, int REFINE>
static void block_fn(unsigned char *out, const unsigned char *rgba, int iw, int w, int h, int nrandom)
{
	encode_one_block(out, rgba, iw, w, h, nrandom, DXT, CD, FAST, REFINE);
}

template <int DXT, int CD, int FAST>
static s2tc_encode_block_func_t pick_refine(int refine)
{
	switch (refine) {
	case REFINE_NEVER: return block_fn<DXT, CD, FAST, REFINE_NEVER>;
	case REFINE_LOOP: return block_fn<DXT, CD, FAST, REFINE_LOOP>;
	default: return block_fn<DXT, CD, FAST, REFINE_ALWAYS>;
	}
}

template <int DXT, int CD>
static s2tc_encode_block_func_t pick_mode(int nrandom, int refine)
{
	if (CD == CD_NORMALMAP || nrandom >= 0) return pick_refine<DXT, CD, 0>(refine);
	return pick_refine<DXT, CD, CD == CD_NORMALMAP ? 0 : 1>(refine);
}

template <int CD>
static s2tc_encode_block_func_t pick_dxt(int dxt, int nrandom, int refine)
{
	switch (dxt) {
	case DXT1: return pick_mode<DXT1, CD>(nrandom, refine);
	case DXT3: return pick_mode<DXT3, CD>(nrandom, refine);
	default: return pick_mode<DXT5, CD>(nrandom, refine);
	}
}

}  // namespace s2tc

using namespace s2tc;

extern "C" {

void tx_compress_dxtn(GLint srccomps, GLint width, GLint height, const GLubyte *srcPixData, GLenum destformat,
                      GLubyte *dest, GLint dstRowStride)
{
	const EnvConfig e = read_env();
	int dxt, k;
	switch (destformat) {
	case 0x83F0: /* GL_COMPRESSED_RGB_S3TC_DXT1_EXT */
	case 0x83F1: /* GL_COMPRESSED_RGBA_S3TC_DXT1_EXT */
		dxt = s2tc::DXT1;
		k = 2;
		break;
	case 0x83F2: dxt = s2tc::DXT3; k = 4; break;
	case 0x83F3: dxt = s2tc::DXT5; k = 4; break;
	default:
		fprintf(stderr, "libdxtn: Bad dstFormat %d in tx_compress_dxtn\n", destformat);
		return;
	}
	if (width <= 0 || height <= 0) return;
	CtxGuard g;
	if (!g.c) {
		fprintf(stderr, "s2tc: cannot create a HIP stream\n");
		return;
	}
	if (int r = encode_host_image(g.c, srccomps, width, height, srcPixData, dxt, e.cd, e.nrandom, e.refine, e.dither,
	                              e.seed)) {
		fprintf(stderr, "s2tc: tx_compress_dxtn failed (%d): %s\n", r, s2tc_amd_last_error());
		return;
	}
	// destination layout with dstRowDiff (s2tc_libtxc_dxtn.cpp:243, :261, :279)
	const int bs = block_bytes(dxt);
	const int bw = (width + 3) / 4, bh = (height + 3) / 4;
	const long rowdiff = dstRowStride >= width * k ? (long)dstRowStride - (long)((width + 3) & ~3) * k : 0;
	const size_t rowbytes = (size_t)bw * bs;
	const uint8_t *hout = g.c->host_out.data();
	auto copy_rows = [&](int lo, int hi) {
		for (int by = lo; by < hi; ++by)
			memcpy(dest + ((long)rowbytes + rowdiff) * by, hout + (size_t)by * rowbytes, rowbytes);
	};
	const unsigned hw = std::thread::hardware_concurrency();
	const unsigned nt = std::max(1u, std::min(kUploadThreads, hw ? hw : 1u));
	if (rowdiff < 0 || nt == 1 || rowbytes * (size_t)bh < (2u << 20)) {
		// overlapping rows (a negative row gap) must be written in order
		copy_rows(0, bh);
		return;
	}
	// disjoint rows: several host threads, contiguous row ranges each
	std::vector<std::thread> workers;
	const int per = (bh + (int)nt - 1) / (int)nt;
	for (unsigned t = 1; t < nt; ++t) {
		const int lo = std::min(bh, per * (int)t), hi = std::min(bh, lo + per);
		if (hi > lo) workers.emplace_back(copy_rows, lo, hi);
	}
	copy_rows(0, std::min(bh, per));
	for (auto &w : workers) w.join();
}

void rgb565_image(unsigned char *out, const unsigned char *rgba, int w, int h, int srccomps, int alphabits,
                  enum DitherMode dither)
{
	if (w <= 0 || h <= 0) return;
	CtxGuard g;
	Ctx *c = g.c;
	if (!c) {
		fprintf(stderr, "s2tc: cannot create a HIP stream\n");
		return;
	}
	const int sc = srccomps == 3 ? 3 : 4;
	const size_t n = (size_t)w * h;
	int r = c->src.ensure(n * sc);
	if (!r) r = c->q.ensure(n * 4);
	if (!r) r = c->scratch.ensure(rgb565_scratch_bytes(w, h));
	if (!r) r = upload(c, c->src.p, rgba, n * sc);
	if (!r) r = launch_rgb565(c->src.p, sc, w, h, alphabits, (int)dither, c->q.p, c->scratch.p, c->stream);
	if (!r) r = hip_check(hipMemcpyAsync(out, c->q.p, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
	if (!r) r = hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
	if (r) fprintf(stderr, "s2tc: rgb565_image failed (%d): %s\n", r, s2tc_amd_last_error());
}

s2tc_encode_block_func_t s2tc_encode_block_func(enum DxtMode dxt, ColorDistMode cd, int nrandom,
                                                enum RefinementMode refine)
{
	switch ((int)cd) {
	case CD_RGB: return pick_dxt<CD_RGB>(dxt, nrandom, refine);
	case CD_YUV: return pick_dxt<CD_YUV>(dxt, nrandom, refine);
	case CD_SRGB: return pick_dxt<CD_SRGB>(dxt, nrandom, refine);
	case CD_SRGB_MIXED: return pick_dxt<CD_SRGB_MIXED>(dxt, nrandom, refine);
	case CD_AVG: return pick_dxt<CD_AVG>(dxt, nrandom, refine);
	case CD_W0AVG: return pick_dxt<CD_W0AVG>(dxt, nrandom, refine);
	case CD_NORMALMAP: return pick_dxt<CD_NORMALMAP>(dxt, nrandom, refine);
	default: return pick_dxt<CD_WAVG>(dxt, nrandom, refine);
	}
}

void s2tc_amd_set_block_index(unsigned int index) { t_block_index = index; }

// ---- S2TC decoders (s2tc_libtxc_dxtn.cpp:35-140) ----

static inline void expand565(unsigned c, GLubyte *t)
{
	unsigned r = (c >> 11) & 0x1f, g = (c >> 5) & 0x3f, b = c & 0x1f;
	t[0] = (GLubyte)((r << 3) | (r >> 2));
	t[1] = (GLubyte)((g << 2) | (g >> 4));
	t[2] = (GLubyte)((b << 3) | (b >> 2));
}

static inline const GLubyte *block_at(GLint stride, const GLubyte *pix, GLint i, GLint j, int bs)
{
	return pix + ((size_t)((stride + 3) >> 2) * (size_t)(j >> 2) + (size_t)(i >> 2)) * bs;
}

// colour of an S2TC block: 0 -> c0, 1 -> c1, 2/3 -> checkerboard of c0/c1;
// with dxt1_black, index 3 with c1 >= c0 is black (transparent for RGBA).
static inline unsigned s2tc_color(const GLubyte *cb, GLint i, GLint j, bool dxt1, bool &black)
{
	const unsigned c0 = cb[0] | (cb[1] << 8), c1 = cb[2] | (cb[3] << 8);
	const int idx = (cb[4 + (j & 3)] >> (2 * (i & 3))) & 3;
	black = false;
	if (idx == 0) return c0;
	if (idx == 1) return c1;
	if (dxt1 && idx == 3 && c1 >= c0) {
		black = true;
		return 0;
	}
	return ((i ^ j) & 1) ? c1 : c0;
}

void fetch_2d_texel_rgb_dxt1(GLint srcRowStride, const GLubyte *pixdata, GLint i, GLint j, GLvoid *texel)
{
	GLubyte *t = (GLubyte *)texel;
	bool black;
	expand565(s2tc_color(block_at(srcRowStride, pixdata, i, j, 8), i, j, true, black), t);
	t[3] = 255;
}

void fetch_2d_texel_rgba_dxt1(GLint srcRowStride, const GLubyte *pixdata, GLint i, GLint j, GLvoid *texel)
{
	GLubyte *t = (GLubyte *)texel;
	bool black;
	expand565(s2tc_color(block_at(srcRowStride, pixdata, i, j, 8), i, j, true, black), t);
	t[3] = black ? 0 : 255;
}

void fetch_2d_texel_rgba_dxt3(GLint srcRowStride, const GLubyte *pixdata, GLint i, GLint j, GLvoid *texel)
{
	GLubyte *t = (GLubyte *)texel;
	const GLubyte *blk = block_at(srcRowStride, pixdata, i, j, 16);
	bool black;
	expand565(s2tc_color(blk + 8, i, j, false, black), t);
	const int a = (blk[(j & 3) * 2 + ((i & 3) >> 1)] >> (4 * (i & 1))) & 0x0f;
	t[3] = (GLubyte)(a | (a << 4));
}

void fetch_2d_texel_rgba_dxt5(GLint srcRowStride, const GLubyte *pixdata, GLint i, GLint j, GLvoid *texel)
{
	GLubyte *t = (GLubyte *)texel;
	const GLubyte *blk = block_at(srcRowStride, pixdata, i, j, 16);
	bool black;
	expand565(s2tc_color(blk + 8, i, j, false, black), t);
	unsigned a = blk[0];
	const unsigned a1 = blk[1];
	uint64_t bits = 0;
	for (int q = 0; q < 6; ++q) bits |= (uint64_t)blk[2 + q] << (8 * q);
	const int ab = (int)((bits >> (3 * ((j & 3) * 4 + (i & 3)))) & 7);
	// index 6 is 0 and 7 is 255 in 6-value mode (a1 >= a0); otherwise 6 falls
	// through to 7's test and both fall through to the checkerboard (:130-137)
	if (ab == 1) a = a1;
	else if (ab == 6 && a1 >= a) a = 0;
	else if ((ab == 6 || ab == 7) && a1 >= a) a = 255;
	else if (ab != 0 && ((i ^ j) & 1)) a = a1;
	t[3] = (GLubyte)a;
}

}  // extern "C"
