// decode.hip - S2TC block decoding and the total colour distance of an
// encoding, on the device.
//
// k_decode follows fetch_2d_texel_* (s2tc_libtxc_dxtn.cpp:35-140) for every
// texel of an image: index 0 / 1 take c0 / c1, indices 2 and 3 take the
// checkerboard of c0 and c1 (i ^ j odd -> c1), except DXT1 index 3 with
// c1 >= c0, which is black (alpha 0 for RGBA); 5:6:5 expands as (v << 3) |
// (v >> 2) and (v << 2) | (v >> 4); DXT3 alpha a | a << 4; DXT5 alpha in
// the 6-value mode (a1 >= a0: 6 -> 0, 7 -> 255, 6 falling through to 7's
// test otherwise) else the checkerboard.  One thread per texel.
//
// The colour-distance kernels live with the encoders (kernels_cd.hip),
// one per distance; s2tc_amd_colordist_error dispatches to them.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "s2tc_internal.h"

namespace s2tc {

__device__ __forceinline__ uint32_t expand565(uint32_t c)
{
	const uint32_t r = (c >> 11) & 31u, g = (c >> 5) & 63u, b = c & 31u;
	return ((r << 3) | (r >> 2)) | (((g << 2) | (g >> 4)) << 8) | (((b << 3) | (b >> 2)) << 16);
}

// DXT5 alpha palette slot for (index ab, texel odd, 6-value mode) at bit
// 2 * (ab * 4 + odd * 2 + six): 0 = a0, 1 = a1, 2 = 0, 3 = 255.
constexpr uint64_t alpha_slots()
{
	uint64_t t = 0;
	for (unsigned ab = 0; ab < 8; ++ab)
		for (unsigned odd = 0; odd < 2; ++odd)
			for (unsigned six = 0; six < 2; ++six) {
				unsigned slot = ab == 0 ? 0u : ab == 1 ? 1u : odd ? 1u : 0u;
				if (six && ab == 6) slot = 2;
				if (six && ab == 7) slot = 3;
				t |= (uint64_t)slot << (2 * (ab * 4 + odd * 2 + six));
			}
	return t;
}
constexpr uint64_t kAlphaSlots = alpha_slots();

// Colour slot for (index b, texel odd, DXT1 block with c1 >= c0) at bit
// 2 * (b * 4 + odd * 2 + black): 0 = c0, 1 = c1, 2 = black.
constexpr uint32_t colour_slots()
{
	uint32_t t = 0;
	for (unsigned b = 0; b < 4; ++b)
		for (unsigned odd = 0; odd < 2; ++odd)
			for (unsigned black = 0; black < 2; ++black) {
				unsigned slot = b == 0 ? 0u : b == 1 ? 1u : odd ? 1u : 0u;
				if (black && b == 3) slot = 2;
				t |= slot << (2 * (b * 4 + odd * 2 + black));
			}
	return t;
}
constexpr uint32_t kColourSlots = colour_slots();

__global__ __launch_bounds__(256) void k_decode(const uint8_t *__restrict__ blocks, unsigned fmt, int width,
                                                int height, uint32_t *__restrict__ rgba)
{
	const int i = (int)(blockIdx.x * 64 + threadIdx.x), j = (int)(blockIdx.y * 4 + threadIdx.y);
	if (i >= width || j >= height) return;
	const int bpb = (fmt == 0x83F0u || fmt == 0x83F1u) ? 8 : 16;
	const uint8_t *blk = blocks + ((size_t)((width + 3) >> 2) * (size_t)(j >> 2) + (size_t)(i >> 2)) * bpb;
	const uint8_t *cb = bpb == 8 ? blk : blk + 8;
	const uint32_t c0 = cb[0] | (cb[1] << 8);
	const uint32_t c1 = cb[2] | (cb[3] << 8);
	const uint32_t b = (uint32_t)(cb[4 + (j & 3)] >> (2 * (i & 3))) & 3u;
	const bool odd = ((i ^ j) & 1) != 0;
	// colour slot (0: c0, 1: c1, 2: black) from (index, odd, DXT1 with
	// c1 >= c0), then the 16-bit field of that slot: shifts only
	const uint32_t black = (bpb == 8 && c1 >= c0) ? 1u : 0u;
	const uint32_t slot = (uint32_t)(kColourSlots >> (2u * (b * 4u + (odd ? 2u : 0u) + black))) & 3u;
	const uint32_t c = (uint32_t)(((uint64_t)c0 | ((uint64_t)c1 << 16)) >> (16u * slot)) & 0xffffu;
	uint32_t a = (fmt == 0x83F1u && slot == 2u) ? 0u : 255u;
	if (fmt == 0x83F2u) {
		const uint32_t q = (blk[(j & 3) * 2 + ((i & 3) >> 1)] >> (4 * (i & 1))) & 15u;
		a = q | (q << 4);
	} else if (fmt == 0x83F3u) {
		// the texel's 3-bit index from the two bytes that hold it (the
		// second may be colour byte 8, whose bits are then not used)
		const uint32_t bitpos = 3u * (uint32_t)((j & 3) * 4 + (i & 3));
		const uint32_t q = bitpos >> 3, sh = bitpos & 7u;
		const uint32_t two = (uint32_t)blk[2 + q] | ((uint32_t)blk[3 + q] << 8);
		const uint32_t ab = (two >> sh) & 7u;
		const uint32_t a0 = blk[0], a1 = blk[1];
		const uint32_t six = a1 >= a0 ? 1u : 0u;  // 6-value mode: indices 6 and 7 are 0 and 255
		// palette slot (0: a0, 1: a1, 2: 0, 3: 255) from (index, odd, six),
		// then the byte of that slot: shifts only, no per-lane branches
		const uint32_t slot = (uint32_t)(kAlphaSlots >> (2u * (ab * 4u + (odd ? 2u : 0u) + six))) & 3u;
		a = ((a0 | (a1 << 8) | 0xff000000u) >> (8u * slot)) & 0xffu;
	}
	rgba[(size_t)j * (size_t)width + (size_t)i] = expand565(c) | (a << 24);
}

}  // namespace s2tc

using namespace s2tc;

extern "C" int s2tc_amd_decode_image(const void *d_blocks, unsigned int format, int width, int height, void *d_rgba,
                                     void *stream)
{
	if (!d_blocks || !d_rgba || width <= 0 || height <= 0)
		return set_error(S2TC_E_ARG, "decode: bad arguments");
	if (format < 0x83F0u || format > 0x83F3u) return set_error(S2TC_E_ARG, "decode: bad format 0x%x", format);
	if (int r = ensure_device()) return r;
	const int by = (height + 3) / 4;
	if (by > 65535) return set_error(S2TC_E_RANGE, "decode: image too tall");
	dim3 grid((unsigned)((width + 63) / 64), (unsigned)by), block(64, 4);
	k_decode<<<grid, block, 0, (hipStream_t)stream>>>((const uint8_t *)d_blocks, format, width, height,
	                                                   (uint32_t *)d_rgba);
	return hip_check(hipGetLastError(), "launch k_decode");
}

extern "C" int s2tc_amd_colordist_error(const void *d_src, int src_kind, int width, int height, const void *d_blocks,
                                        int dxt, int cd, long long *total, void *stream)
{
	if (!d_src || !d_blocks || !total || width <= 0 || height <= 0)
		return set_error(S2TC_E_ARG, "colordist_error: bad arguments");
	if (src_kind != S2TC_SRC_QUANTIZED && src_kind != S2TC_SRC_RGBA8)
		return set_error(S2TC_E_ARG, "colordist_error: bad source kind %d", src_kind);
	int refine = REFINE_ALWAYS;
	normalize_modes(dxt, cd, refine);
	if (int r = ensure_device()) return r;
	const KernelTable &kt = kernel_table(cd);
	const void *k = kt.error[dxt][src_kind == S2TC_SRC_RGBA8 ? 1 : 0];
	if (!k) return set_error(S2TC_E_ARG, "colordist_error: no kernel for distance %d", cd);
	const int bw = (width + 3) / 4, bh = (height + 3) / 4;
	if (bh > 4 * 65535) return set_error(S2TC_E_RANGE, "colordist_error: image too tall");
	hipStream_t st = (hipStream_t)stream;
	long long *d_sum = nullptr;
	int rc = hip_check(hipMalloc((void **)&d_sum, sizeof(long long)), "hipMalloc");
	if (!rc) rc = hip_check(hipMemsetAsync(d_sum, 0, sizeof(long long), st), "hipMemsetAsync");
	if (!rc) {
		void *args[] = {(void *)&d_src, (void *)&width, (void *)&height, (void *)&d_blocks, (void *)&d_sum};
		dim3 grid((unsigned)((bw + 63) / 64), (unsigned)((bh + 3) / 4)), block(64, 4);
		rc = hip_check(hipLaunchKernel(k, grid, block, args, 0, st), "launch colour-distance kernel");
	}
	if (!rc) rc = hip_check(hipMemcpyAsync(total, d_sum, sizeof(long long), hipMemcpyDeviceToHost, st), "D2H");
	if (!rc) rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
	if (d_sum) (void)hipFree(d_sum);
	return rc;
}
