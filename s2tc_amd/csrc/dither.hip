// dither.hip - rgb565_image on the GPU (s2tc_algorithm.cpp:1196-1465).
//
// DITHER_NONE is a per-texel shift (quantize_none).
//
// DITHER_SIMPLE carries one running error per channel through the whole
// image in row-major order (:1307-1349).  The error is a small finite state:
// diffuse() with shift s keeps d in [-(2^s - 1), 2^s - 1] (15 states for
// shift 3, 7 for shift 2, 31 for shift 4), and diffuse1() keeps
// e = d + 127 = (127 + sum of inputs) mod 255.  The chain is therefore a
// scan over transition maps, all channels of a call in each launch
// (blockIdx.y = channel):
//   1. chunk maps:   for every chunk of L texels and every start state, the
//                    end state (one thread per (chunk, state));
//   2. super maps:   compositions of K consecutive chunk maps (staged in LDS);
//   3. super starts: one thread walks the super maps from the initial state;
//   4. emit:         one workgroup per super chunk: thread k walks the chunk
//                    maps before its chunk (in LDS) to its start state, then
//                    re-runs its texels and writes the quantized values.
// L (16..256) and K (16..256) grow with the image so that small images get
// short serial walks and many threads.  The mod-255 alpha channel uses sums
// instead of maps.
//
// DITHER_FLOYDSTEINBERG (:1350-1412) diffuses a 12-bit error to the right
// (7/16) and to the row below (3, 5 and the rest of the remainder).  Texel
// (x, y) reads the accumulated error A_y[x+1] that row y-1 finalises when it
// processes texel x+1, so rows form a wavefront with a skew of two texels:
// thread r of one workgroup handles row y0 + r and at step t texel
// x = t - 2r.  Each row keeps a window of its not yet final downward errors
// in registers and hands A_{y+1}[x] to the next row through LDS (double
// buffered by step parity); the last row of a band hands them to the first
// row of the next band through a global carry row.  The reference reuses one
// error buffer for the colour and the alpha pass and does not clear the
// part the alpha pass reads for its first row: that row starts from the
// red-channel errors the colour pass left there (see k_dither_fs).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "s2tc_internal.h"

namespace s2tc {

#define DITHER_S 32    // map stride: states per chunk map (at most 31 used)

struct DChan {
	int off;     // byte offset of the channel within a source texel
	int kind;    // 0: diffuse(shift), 1: diffuse1 (mod 255)
	int shift;
	int nstates; // 2 * (2^shift - 1) + 1 for kind 0, 1 for kind 1
};

__device__ __forceinline__ int diffuse_step(int d, int src, int shift, int &ret)
{
	const int maxval = (1 << (8 - shift)) - 1;
	const int s = src + d;
	int r = s >> shift;
	r = r < 0 ? 0 : (r > maxval ? maxval : r);
	ret = r;
	return s - ((r << shift) | (r >> (8 - 2 * shift)));
}

// The SIMPLE channels of one call and the chunking.
struct DChans {
	DChan ch[4];
	int nch;
	int L, K;             // texels per chunk, chunks per super chunk
	long long nchunks, nsuper;
	long long stride;     // scratch bytes per channel
};

// per-channel scratch: chunk maps (nchunks x 32), super maps (nsuper x 32),
// super starts (nsuper)
__device__ __forceinline__ uint8_t *chan_maps(uint8_t *scr, const DChans &cs, int c) { return scr + c * cs.stride; }
__device__ __forceinline__ uint8_t *chan_smaps(uint8_t *scr, const DChans &cs, int c)
{
	return scr + c * cs.stride + cs.nchunks * DITHER_S;
}
__device__ __forceinline__ uint8_t *chan_sstart(uint8_t *scr, const DChans &cs, int c)
{
	return scr + c * cs.stride + (cs.nchunks + cs.nsuper) * DITHER_S;
}

// 1. end state of every (chunk, start state); the texel loads of a chunk are
// independent of the state, eight at a time ahead of the chain
__global__ __launch_bounds__(256) void k_dither_chunk_maps(const uint8_t *__restrict__ src, int srccomps, long long n,
                                                           DChans cs, uint8_t *__restrict__ scr)
{
	const DChan ch = cs.ch[blockIdx.y];
	const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const long long c = tid >> 5;
	const int s0 = (int)(tid & 31);
	const int S = ch.nstates;
	if (c >= cs.nchunks || s0 >= S) return;
	uint8_t *maps = chan_maps(scr, cs, blockIdx.y);
	const long long p0 = c * cs.L, p1 = min(p0 + cs.L, n);
	const uint8_t *sp = src + ch.off;
	if (ch.kind == 1) {
		int acc = 0;
		for (long long p = p0; p < p1; ++p) acc += sp[p * srccomps];
		maps[c * DITHER_S] = (uint8_t)(acc % 255);
		return;
	}
	const int half = (S - 1) / 2;
	int d = s0 - half, ret;
	long long p = p0;
	for (; p + 8 <= p1; p += 8) {
		int v[8];
#pragma unroll
		for (int u = 0; u < 8; ++u) v[u] = sp[(p + u) * srccomps];
#pragma unroll
		for (int u = 0; u < 8; ++u) d = diffuse_step(d, v[u], ch.shift, ret);
	}
	for (; p < p1; ++p) d = diffuse_step(d, sp[p * srccomps], ch.shift, ret);
	maps[c * DITHER_S + s0] = (uint8_t)(d + half);
}

// 2. composition of the K chunk maps of every super chunk; a workgroup of
// 128 threads stages the maps of 4 super chunks in LDS, thread (g, state)
// walks super chunk g
__global__ __launch_bounds__(128) void k_dither_super_maps(DChans cs, uint8_t *__restrict__ scr)
{
	extern __shared__ uint8_t lds[];
	const DChan ch = cs.ch[blockIdx.y];
	const uint8_t *maps = chan_maps(scr, cs, blockIdx.y);
	uint8_t *smaps = chan_smaps(scr, cs, blockIdx.y);
	const long long sc0 = (long long)blockIdx.x * 4;
	const long long c0 = sc0 * cs.K, c1 = min(c0 + 4LL * cs.K, cs.nchunks);
	const long long bytes = (c1 - c0) * DITHER_S;
	const uint32_t *g = (const uint32_t *)(maps + c0 * DITHER_S);
	for (long long i = threadIdx.x; i < bytes / 4; i += blockDim.x) ((uint32_t *)lds)[i] = g[i];
	__syncthreads();
	const int gi = threadIdx.x >> 5, s0 = threadIdx.x & 31;
	const long long sc = sc0 + gi;
	if (sc >= cs.nsuper) return;
	const int S = ch.nstates;
	const long long a = (long long)gi * cs.K, b = min(a + cs.K, c1 - c0);
	if (ch.kind == 1) {
		if (s0) return;
		int acc = 0;
		for (long long c = a; c < b; ++c) acc += lds[c * DITHER_S];
		smaps[sc * DITHER_S] = (uint8_t)(acc % 255);
		return;
	}
	if (s0 >= S) return;
	int st = s0;
	for (long long c = a; c < b; ++c) st = lds[c * DITHER_S + st];
	smaps[sc * DITHER_S + s0] = (uint8_t)st;
}

// 3. start state of every super chunk: one thread per channel walks the
// super maps (staged in LDS when they fit).  With a carry (a band of a larger
// image), the walk starts from carry[channel] when carry_in is set and leaves
// the end state there for the next band.
__global__ __launch_bounds__(256) void k_dither_super_starts(DChans cs, uint8_t *__restrict__ scr, int staged,
                                                             uint8_t *__restrict__ carry, int carry_in)
{
	extern __shared__ uint8_t lds[];
	const DChan ch = cs.ch[blockIdx.y];
	const uint8_t *smaps = chan_smaps(scr, cs, blockIdx.y);
	uint8_t *sstart = chan_sstart(scr, cs, blockIdx.y);
	const uint8_t *m = smaps;
	if (staged) {
		const long long words = cs.nsuper * DITHER_S / 4;
		for (long long i = threadIdx.x; i < words; i += blockDim.x) ((uint32_t *)lds)[i] = ((const uint32_t *)smaps)[i];
		__syncthreads();
		m = lds;
	}
	if (threadIdx.x != 0) return;
	const int S = ch.nstates;
	int st = ch.kind == 1 ? 127 : (S - 1) / 2;  // d = 0 (e = 127 for the mod-255 channel)
	if (carry && carry_in) st = carry[blockIdx.y];
	for (long long sc = 0; sc < cs.nsuper; ++sc) {
		sstart[sc] = (uint8_t)st;
		st = ch.kind == 1 ? (st + m[sc * DITHER_S]) % 255 : m[sc * DITHER_S + st];
	}
	if (carry) carry[blockIdx.y] = (uint8_t)st;
}

// 4. one workgroup (K threads) per super chunk: thread k finds the start
// state of chunk k of the super chunk from the staged chunk maps, then
// re-runs the chunk's texels and writes the channel
__global__ __launch_bounds__(256) void k_dither_emit(const uint8_t *__restrict__ src, int srccomps, long long n,
                                                     DChans cs, uint8_t *__restrict__ scr, uint8_t *__restrict__ out)
{
	extern __shared__ uint8_t lds[];
	const DChan ch = cs.ch[blockIdx.y];
	const uint8_t *maps = chan_maps(scr, cs, blockIdx.y);
	const uint8_t *sstart = chan_sstart(scr, cs, blockIdx.y);
	const long long sc = blockIdx.x;
	const long long c0 = sc * cs.K, c1 = min(c0 + cs.K, cs.nchunks);
	const long long words = (c1 - c0) * DITHER_S / 4;
	const uint32_t *g = (const uint32_t *)(maps + c0 * DITHER_S);
	for (long long i = threadIdx.x; i < words; i += blockDim.x) ((uint32_t *)lds)[i] = g[i];
	__syncthreads();
	const long long c = c0 + threadIdx.x;
	if (c >= c1) return;
	int st = sstart[sc];
	for (int k = 0; k < (int)threadIdx.x; ++k)
		st = ch.kind == 1 ? (st + lds[k * DITHER_S]) % 255 : lds[k * DITHER_S + st];
	const long long p0 = c * cs.L, p1 = min(p0 + cs.L, n);
	const uint8_t *sp = src + ch.off;
	if (ch.kind == 1) {
		int e = st;
		for (long long p = p0; p < p1; ++p) {
			const int v = sp[p * srccomps];
			out[p * 4 + ch.off] = (uint8_t)(e + v >= 255 ? 1 : 0);  // d + src >= 128 with e = d + 127
			e = (e + v) % 255;
		}
		return;
	}
	const int half = (ch.nstates - 1) / 2;
	int d = st - half, ret;
	long long p = p0;
	for (; p + 8 <= p1; p += 8) {
		int v[8];
#pragma unroll
		for (int u = 0; u < 8; ++u) v[u] = sp[(p + u) * srccomps];
#pragma unroll
		for (int u = 0; u < 8; ++u) {
			d = diffuse_step(d, v[u], ch.shift, ret);
			out[(p + u) * 4 + ch.off] = (uint8_t)ret;
		}
	}
	for (; p < p1; ++p) {
		d = diffuse_step(d, sp[p * srccomps], ch.shift, ret);
		out[p * 4 + ch.off] = (uint8_t)ret;
	}
}

// DITHER_NONE (:1269-1306) and the constant / copied alpha of every mode
__global__ void k_quantize_none(const uint8_t *__restrict__ src, int srccomps, long long n, int alphabits,
                                int rgb, int alpha, uint8_t *__restrict__ out)
{
	const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n) return;
	const uint8_t *s = src + p * srccomps;
	if (rgb) {
		out[p * 4 + 0] = s[0] >> 3;
		out[p * 4 + 1] = s[1] >> 2;
		out[p * 4 + 2] = s[2] >> 3;
	}
	if (alpha) {
		if (srccomps == 4) {
			const int a = s[3];
			out[p * 4 + 3] = (uint8_t)(alphabits == 1 ? a >> 7 : alphabits == 8 ? a : a >> (8 - alphabits));
		} else {
			out[p * 4 + 3] = (uint8_t)((1 << alphabits) - 1);
		}
	}
}

// e7, e3, e5, e1 of floyd/floyd1 (:1227-1233), C division (toward zero)
__device__ __forceinline__ void fs_split(int err, int &e7, int &e3, int &e5, int &e1)
{
	e7 = (err * 7 + 8) / 16;
	err -= e7;
	e3 = (err * 3 + 4) / 9;
	err -= e3;
	e5 = (err * 5 + 3) / 6;
	err -= e5;
	e1 = err;
}

// ---- Floyd-Steinberg ----
// Rows per band R (the workgroup size) is a template parameter: 256, 512 or
// 1024, the smallest that covers min(h, 1024).  The wavefront takes
// ceil(h / R) * (w + 2R - 1) steps, each a dependent chain of about fifty
// instructions plus an LDS hand-off and a barrier, so tall bands cut the
// step count (a 4096^2 image: 24,572 steps at R = 1024 against 73,712 at 256).

struct FsChan {
	int off;    // byte offset of the channel in a source texel
	int kind;   // 0: floyd(shift), 1: floyd1
	int shift;
};

__device__ __forceinline__ int fs_load(const int *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fs_store(int *p, int v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One Floyd-Steinberg pass per workgroup: workgroup k runs channel k
// (c0, c1, c2) with its own carry row at carry + k * (w + 2); the channels
// of the reference's colour pass are independent (separate error rows,
// s2tc_algorithm.cpp:1356-1372), so they run side by side on different CUs.
// init0 (w+2 ints, may be null) is A_0 of workgroup 0's channel, the errors
// its first row starts from (zero otherwise).  rleft (w+2 ints, may be null)
// receives what the reference's buffer holds in the first w+2 ints after the
// pass, for workgroup 0's channel: the accumulated errors the last row read
// (thisrow, when h-1 is even) or that row's downward errors (downrow, when
// h-1 is odd).  The quantized step divides by maxval (15, 31 or 63) with one
// high multiply: floor(v / d) = mul_hi(v, floor(2^32 / d) + 1) for v < 2^20
// and odd d < 2^12, since the product overshoots v / d by less than
// 2^-12 < 1 / d.
template <int FS_R>
__global__ __launch_bounds__(FS_R) void k_dither_fs(const uint8_t *__restrict__ src, int srccomps, int w, int h,
                                                    FsChan c0, FsChan c1, FsChan c2, const int *__restrict__ init0,
                                                    int *__restrict__ carry_all, int *__restrict__ rleft_all,
                                                    uint8_t *__restrict__ out)
{
	__shared__ int slot[FS_R][2];
	const int k = (int)blockIdx.x;
	const FsChan ch = k == 0 ? c0 : k == 1 ? c1 : c2;
	const int r = (int)threadIdx.x;
	const int pw = w + 2;
	int *carry = carry_all + (size_t)k * pw;
	const int *init = k == 0 ? init0 : nullptr;
	int *rleft = k == 0 ? rleft_all : nullptr;
	const int maxval = (1 << (8 - ch.shift)) - 1;
	const uint32_t rec = 0xffffffffu / (uint32_t)maxval + 1u;
	for (int j = r; j < pw; j += FS_R) fs_store(&carry[j], init ? init[j] : 0);
	__syncthreads();
	for (int y0 = 0; y0 < h; y0 += FS_R) {
		const int rows = min(FS_R, h - y0);
		const int y = y0 + r;
		const bool active = r < rows;
		const uint8_t *srow = src + (size_t)(active ? y : 0) * (size_t)w * (size_t)srccomps + ch.off;
		uint8_t *orow = out + (size_t)(active ? y : 0) * (size_t)w * 4 + ch.off;
		int right = 0, win1 = 0, win2 = 0;
		const int steps = w + 2 * (rows - 1) + 1;
		for (int t = 0; t < steps; ++t) {
			const int x = t - 2 * r;
			if (active && x >= 0 && x <= w) {
				int fin;
				if (x < w) {
					const int above = r == 0 ? fs_load(&carry[x + 1]) : slot[r - 1][(t - 1) & 1];
					const int acc = above + right;
					const int sv = srow[x * srccomps];
					const int v = ((sv << 4) | (sv >> 4)) + acc;
					int q, err;
					if (ch.kind == 1) {
						q = v >= 2048 ? 1 : 0;
						err = v - (q ? 4095 : 0);
					} else {
						q = v >> (ch.shift + 4);
						q = q < 0 ? 0 : (q > maxval ? maxval : q);
						err = v - (int)__umulhi((uint32_t)(q * 4095), rec);
					}
					orow[x * 4] = (uint8_t)q;
					int e7, e3, e5, e1;
					fs_split(err, e7, e3, e5, e1);
					right = e7;
					fin = win1 + e3;
					win1 = win2 + e5;
					win2 = e1;
					if (rleft && y == h - 1 && ((h - 1) & 1) == 0) rleft[x + 1] = acc;
				} else {
					fin = win1;
				}
				if (r + 1 < rows) slot[r][t & 1] = fin;  // A_{y+1}[x] for row y+1
				else fs_store(&carry[x], fin);
				if (rleft && y == h - 1 && ((h - 1) & 1) == 1) rleft[x] = fin;
			}
			__syncthreads();
		}
	}
}

template <int FS_R>
static void fs_pass(int nch, const uint8_t *src, int srccomps, int w, int h, FsChan c0, FsChan c1, FsChan c2,
                    const int *init0, int *carry, int *rleft, uint8_t *out, hipStream_t st)
{
	k_dither_fs<FS_R><<<nch, FS_R, 0, st>>>(src, srccomps, w, h, c0, c1, c2, init0, carry, rleft, out);
}

static int launch_rgb565_fs(const uint8_t *src, int srccomps, int w, int h, int alphabits, uint8_t *out,
                            void *d_scratch, hipStream_t st)
{
	int *carry = (int *)d_scratch;          // 3 * (w + 2)
	int *rleft = carry + 3 * (w + 2);        // w + 2
	const bool alpha_fs = srccomps == 4 && (alphabits == 1 || alphabits == 4);
	const FsChan cr{0, 0, 3}, cg{1, 0, 2}, cb{2, 0, 3};
	const int R = h <= 256 ? 256 : h <= 512 ? 512 : 1024;
	int *rl = alpha_fs ? rleft : nullptr;
	// r, g and b: three workgroups, one channel each
	if (R == 256) fs_pass<256>(3, src, srccomps, w, h, cr, cg, cb, nullptr, carry, rl, out, st);
	else if (R == 512) fs_pass<512>(3, src, srccomps, w, h, cr, cg, cb, nullptr, carry, rl, out, st);
	else fs_pass<1024>(3, src, srccomps, w, h, cr, cg, cb, nullptr, carry, rl, out, st);
	if (int r = hip_check(hipGetLastError(), "dither FS colour kernel")) return r;
	if (alpha_fs) {
		// alpha starts from what the r pass left in the shared buffer
		const FsChan ca{3, alphabits == 1 ? 1 : 0, 8 - alphabits};
		if (R == 256) fs_pass<256>(1, src, srccomps, w, h, ca, ca, ca, rleft, carry, nullptr, out, st);
		else if (R == 512) fs_pass<512>(1, src, srccomps, w, h, ca, ca, ca, rleft, carry, nullptr, out, st);
		else fs_pass<1024>(1, src, srccomps, w, h, ca, ca, ca, rleft, carry, nullptr, out, st);
		return hip_check(hipGetLastError(), "dither FS alpha kernel");
	}
	const long long n = (long long)w * h;
	k_quantize_none<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(src, srccomps, n, alphabits, 0, 1, out);
	return hip_check(hipGetLastError(), "quantize alpha kernel");
}

// ---- host launchers ----

// Chunk length L and super chunk size K for n texels: powers of two with
// about n / 16384 texels per chunk (16..256) and about sqrt(chunks) chunks per
// super chunk (16..256).
struct SimpleChunking {
	int L, K;
	long long nchunks, nsuper, stride;
};

static constexpr SimpleChunking simple_chunking_of(long long n)
{
	// chunks of at least 32 texels: on a 256x256 image 16-texel chunks made
	// the call 4% slower (more maps to compose, no shorter serial walk)
	int L = 32;
	while (L < 256 && (long long)L * 16384 < n) L <<= 1;
	const long long nchunks = (n + L - 1) / L;
	int K = 16;
	while (K < 256 && (long long)K * K < nchunks) K <<= 1;
	const long long nsuper = (nchunks + K - 1) / K;
	// 4-byte aligned regions (the kernels copy maps as words)
	return SimpleChunking{L, K, nchunks, nsuper, ((nchunks + nsuper) * DITHER_S + nsuper + 255) & ~255LL};
}

static void simple_chunking(long long n, DChans &cs)
{
	const SimpleChunking c = simple_chunking_of(n);
	cs.L = c.L;
	cs.K = c.K;
	cs.nchunks = c.nchunks;
	cs.nsuper = c.nsuper;
	cs.stride = c.stride;
}

// scratch bytes of one dither mode: SIMPLE maps of up to four channels, FS
// carry rows of three channels plus the row handed to the alpha pass
static constexpr size_t simple_scratch_bytes(long long n)
{
	return (size_t)(4 * simple_chunking_of(n > 0 ? n : 1).stride + 256);
}
static constexpr size_t fs_scratch_bytes(int w) { return (size_t)(4 * ((long long)w + 2)) * sizeof(int) + 256; }

size_t rgb565_scratch_bytes(int w, int h)
{
	const size_t simple = simple_scratch_bytes((long long)w * h), fs = fs_scratch_bytes(w);
	return simple > fs ? simple : fs;
}

static int run_simple(const uint8_t *src, int srccomps, long long n, DChans cs, uint8_t *out, uint8_t *scr,
                      uint8_t *carry, int carry_in, hipStream_t st)
{
	simple_chunking(n, cs);
	const unsigned nch = (unsigned)cs.nch;
	k_dither_chunk_maps<<<dim3((unsigned)((cs.nchunks * DITHER_S + 255) / 256), nch), 256, 0, st>>>(src, srccomps, n,
	                                                                                                  cs, scr);
	k_dither_super_maps<<<dim3((unsigned)((cs.nsuper + 3) / 4), nch), 128, (size_t)4 * cs.K * DITHER_S, st>>>(cs, scr);
	const size_t lds = (size_t)(cs.nsuper * DITHER_S);
	const int staged = lds <= 60000;
	k_dither_super_starts<<<dim3(1, nch), 256, staged ? lds : 0, st>>>(cs, scr, staged, carry, carry_in);
	k_dither_emit<<<dim3((unsigned)cs.nsuper, nch), (unsigned)cs.K, (size_t)cs.K * DITHER_S, st>>>(src, srccomps, n, cs,
	                                                                                             scr, out);
	return hip_check(hipGetLastError(), "dither SIMPLE kernels");
}

// rgb565_image with the reference's argument normalisation: srccomps other
// than 3 means 4 (:1455-1463), alphabits other than 1 and 4 means 8
// (:1437-1449), dither other than NONE and FLOYDSTEINBERG means SIMPLE
// (:1419-1431).
int launch_rgb565(const void *d_src, int srccomps, int w, int h, int alphabits, int dither, void *d_out,
                  void *d_scratch, hipStream_t st, void *d_carry, int carry_in)
{
	if (srccomps != 3) srccomps = 4;
	if (alphabits != 1 && alphabits != 4) alphabits = 8;
	if (dither != DITHER_NONE && dither != DITHER_FLOYDSTEINBERG) dither = DITHER_SIMPLE;
	const long long n = (long long)w * h;
	if (n <= 0) return 0;
	const uint8_t *src = (const uint8_t *)d_src;
	uint8_t *out = (uint8_t *)d_out;
	const unsigned grid = (unsigned)((n + 255) / 256);
	if (dither == DITHER_NONE) {
		k_quantize_none<<<grid, 256, 0, st>>>(src, srccomps, n, alphabits, 1, 1, out);
		return hip_check(hipGetLastError(), "quantize kernel");
	}
	if (dither == DITHER_FLOYDSTEINBERG) {
		if (d_carry) return set_error(S2TC_E_ARG, "the Floyd-Steinberg dither does not run in bands");
		return launch_rgb565_fs(src, srccomps, w, h, alphabits, out, d_scratch, st);
	}
	uint8_t *scr = (uint8_t *)d_scratch;
	DChans cs;
	cs.ch[0] = DChan{0, 0, 3, 15};
	cs.ch[1] = DChan{1, 0, 2, 7};
	cs.ch[2] = DChan{2, 0, 3, 15};
	cs.nch = 3;
	if (srccomps == 4 && alphabits == 1) cs.ch[cs.nch++] = DChan{3, 1, 0, 1};
	else if (srccomps == 4 && alphabits == 4) cs.ch[cs.nch++] = DChan{3, 0, 4, 31};
	if (int r = run_simple(src, srccomps, n, cs, out, scr, (uint8_t *)d_carry, carry_in, st)) return r;
	if (cs.nch == 4) return 0;
	k_quantize_none<<<grid, 256, 0, st>>>(src, srccomps, n, alphabits, 0, 1, out);
	return hip_check(hipGetLastError(), "quantize alpha kernel");
}

}  // namespace s2tc
