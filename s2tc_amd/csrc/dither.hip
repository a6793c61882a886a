// dither.hip - rgb565_image on the GPU (s2tc_algorithm.cpp:1196-1465).
//
// DITHER_NONE is a per-texel shift (quantize_none).
//
// DITHER_SIMPLE carries one running error per channel through the whole
// image in row-major order (:1307-1349).  The error is a small finite state:
// diffuse() with shift s keeps d in [-(2^s - 1), 2^s - 1] (15 states for
// shift 3, 7 for shift 2, 31 for shift 4), and diffuse1() keeps
// e = d + 127 = (127 + sum of inputs) mod 255.  The chain is therefore a
// scan over transition maps:
//   1. chunk maps:   for every chunk of L texels and every start state, the
//                    end state (one thread per (chunk, state));
//   2. super maps:   compositions of K consecutive chunk maps;
//   3. super starts: one thread walks the super maps from the initial state;
//   4. chunk starts: one thread per super chunk walks its K chunk maps;
//   5. emit:         one thread per chunk re-runs its texels from the start
//                    state and writes the quantized values.
// The mod-255 alpha channel uses sums instead of maps.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace s2tc {

#define DITHER_L 256   // texels per chunk
#define DITHER_K 256   // chunks per super chunk

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

// 1. end state of every (chunk, start state)
__global__ void k_dither_chunk_maps(const uint8_t *__restrict__ src, int srccomps, long long n, DChan ch,
                                    uint8_t *__restrict__ maps, long long nchunks)
{
	const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const int S = ch.nstates;
	if (tid >= nchunks * S) return;
	const long long c = tid / S;
	const int s0 = (int)(tid - c * S);
	const long long p0 = c * DITHER_L, p1 = min(p0 + DITHER_L, n);
	if (ch.kind == 1) {
		int acc = 0;
		for (long long p = p0; p < p1; ++p) acc += src[p * srccomps + ch.off];
		maps[c] = (uint8_t)(acc % 255);
		return;
	}
	const int half = (S - 1) / 2;
	int d = s0 - half, ret;
	for (long long p = p0; p < p1; ++p) d = diffuse_step(d, src[p * srccomps + ch.off], ch.shift, ret);
	maps[c * S + s0] = (uint8_t)(d + half);
}

// 2. composition of K consecutive chunk maps
__global__ void k_dither_super_maps(const uint8_t *__restrict__ maps, long long nchunks, DChan ch,
                                    uint8_t *__restrict__ smaps, long long nsuper)
{
	const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	const int S = ch.nstates;
	if (tid >= nsuper * S) return;
	const long long sc = tid / S;
	const int s0 = (int)(tid - sc * S);
	const long long c0 = sc * DITHER_K, c1 = min(c0 + DITHER_K, nchunks);
	if (ch.kind == 1) {
		int acc = 0;
		for (long long c = c0; c < c1; ++c) acc += maps[c];
		smaps[sc] = (uint8_t)(acc % 255);
		return;
	}
	int s = s0;
	for (long long c = c0; c < c1; ++c) s = maps[c * S + s];
	smaps[sc * S + s0] = (uint8_t)s;
}

// 3. start state of every super chunk (one thread; super maps staged in LDS)
__global__ void k_dither_super_starts(const uint8_t *__restrict__ smaps, long long nsuper, DChan ch,
                                      uint8_t *__restrict__ sstart)
{
	extern __shared__ uint8_t lds[];
	const int S = ch.nstates;
	const long long total = nsuper * S;
	for (long long i = threadIdx.x; i < total; i += blockDim.x) lds[i] = smaps[i];
	__syncthreads();
	if (threadIdx.x != 0) return;
	int s = ch.kind == 1 ? 127 : (S - 1) / 2;  // d = 0 (e = 127 for the mod-255 channel)
	for (long long sc = 0; sc < nsuper; ++sc) {
		sstart[sc] = (uint8_t)s;
		s = ch.kind == 1 ? (s + lds[sc]) % 255 : lds[sc * S + s];
	}
}

// 4. start state of every chunk
__global__ void k_dither_chunk_starts(const uint8_t *__restrict__ maps, long long nchunks, DChan ch,
                                      const uint8_t *__restrict__ sstart, long long nsuper,
                                      uint8_t *__restrict__ cstart)
{
	const long long sc = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (sc >= nsuper) return;
	const int S = ch.nstates;
	int s = sstart[sc];
	const long long c0 = sc * DITHER_K, c1 = min(c0 + DITHER_K, nchunks);
	for (long long c = c0; c < c1; ++c) {
		cstart[c] = (uint8_t)s;
		s = ch.kind == 1 ? (s + maps[c]) % 255 : maps[c * S + s];
	}
}

// 5. re-run every chunk from its start state and write the channel
__global__ void k_dither_emit(const uint8_t *__restrict__ src, int srccomps, long long n, DChan ch,
                              const uint8_t *__restrict__ cstart, long long nchunks, uint8_t *__restrict__ out)
{
	const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= nchunks) return;
	const long long p0 = c * DITHER_L, p1 = min(p0 + DITHER_L, n);
	if (ch.kind == 1) {
		int e = cstart[c];
		for (long long p = p0; p < p1; ++p) {
			const int v = src[p * srccomps + ch.off];
			out[p * 4 + ch.off] = (uint8_t)(e + v >= 255 ? 1 : 0);  // d + src >= 128 with e = d + 127
			e = (e + v) % 255;
		}
		return;
	}
	const int half = (ch.nstates - 1) / 2;
	int d = cstart[c] - half, ret;
	for (long long p = p0; p < p1; ++p) {
		d = diffuse_step(d, src[p * srccomps + ch.off], ch.shift, ret);
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

}  // namespace s2tc
