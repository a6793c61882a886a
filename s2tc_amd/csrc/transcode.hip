// transcode.hip - S3TC -> S2TC block remap on the device (s2tc_from_s3tc).
//
// s2tc_from_s3tc (s2tc_from_s3tc.cpp:77-190, 254-263) rewrites every block of
// a DXT1/3/5 file so that it only uses what S2TC decoders support: colour
// indices 2 (and 3 outside DXT1's transparent mode) and alpha indices of the
// interpolated values become one of the two endpoints in a checkerboard
// pattern ((x + y) & 1 of the texel), and the endpoints are put in S2TC's
// order (swapping them and flipping the 0/1 indices where needed).  Each
// block is independent: one lane per block, one 8- or 16-byte load and store,
// all index fields of a block handled at once as bit planes of its index
// word(s).  In place is allowed (d_in == d_out).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "s2tc_internal.h"

namespace s2tc {

// low bit of each 2-bit colour index field / of each 3-bit alpha field
constexpr uint32_t kLo2 = 0x55555555u;
constexpr uint64_t kLo3 = 0x249249249249ull;
// the same positions restricted to texels with (x + y) odd
constexpr uint32_t kChk2 = 0x11441144u;
constexpr uint64_t kChk3 = 0x041208041208ull;

// colour block words w0 = c0 | c1 << 16, w1 = 2-bit indices
__device__ __forceinline__ void remap_colour(uint32_t &w0, uint32_t &w1, bool dxt1)
{
	const uint32_t c0 = w0 & 0xffffu, c1 = w0 >> 16;
	const uint32_t lo = w1 & kLo2, hi = (w1 >> 1) & kLo2;
	const bool keep3 = dxt1 && c1 >= c0;
	const bool swap = dxt1 ? c1 < c0 : c1 >= c0;
	uint32_t r;
	if (keep3)   // 0, 1, 3 kept; 2 -> checker
		r = (lo | (hi & kChk2)) | ((hi & lo) << 1);
	else         // 0, 1 kept; 2, 3 -> checker
		r = (lo & ~hi) | (hi & kChk2);
	if (swap) {
		r ^= kLo2;
		w0 = c1 | (c0 << 16);
	}
	w1 = r;
}

// DXT5 alpha half: a = a0 | a1 << 8 | 48 index bits << 16
__device__ __forceinline__ uint64_t remap_alpha(uint64_t a)
{
	const uint32_t a0 = (uint32_t)a & 0xff, a1 = ((uint32_t)a >> 8) & 0xff;
	const uint64_t idx = a >> 16;
	const uint64_t b1 = (idx >> 1) & kLo3, b2 = (idx >> 2) & kLo3;
	const bool six = a1 >= a0;
	const uint64_t t = six ? (b1 ^ b2) : (b1 | b2);   // fields that become the checker value
	uint64_t r = (idx & ~(t * 7)) | (t & kChk3);
	uint64_t e = a & 0xffff;
	if (!six) {
		r ^= kLo3;
		e = a1 | (a0 << 8);
	}
	return e | (r << 16);
}

template <int DXT>
__global__ __launch_bounds__(256) void k_from_s3tc(const void *__restrict__ in, void *__restrict__ out, size_t n)
{
	const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	if constexpr (DXT == 0) {
		uint2 b = ((const uint2 *)in)[i];
		remap_colour(b.x, b.y, true);
		((uint2 *)out)[i] = b;
	} else {
		uint4 b = ((const uint4 *)in)[i];
		remap_colour(b.z, b.w, false);
		if constexpr (DXT == 2) {
			const uint64_t a = remap_alpha((uint64_t)b.x | ((uint64_t)b.y << 32));
			b.x = (uint32_t)a;
			b.y = (uint32_t)(a >> 32);
		}
		((uint4 *)out)[i] = b;
	}
}

}  // namespace s2tc

using namespace s2tc;

extern "C" int s2tc_amd_from_s3tc(const void *d_in, void *d_out, size_t nblocks, int dxt, void *stream)
{
	if (dxt < 0 || dxt > 2) return set_error(S2TC_E_ARG, "from_s3tc: format %d is not DXT1/3/5 (0/1/2)", dxt);
	if (nblocks == 0) return 0;
	if (!d_in || !d_out) return set_error(S2TC_E_ARG, "from_s3tc: null buffer");
	const size_t align = dxt == 0 ? 8 : 16;
	if (((uintptr_t)d_in | (uintptr_t)d_out) % align)
		return set_error(S2TC_E_ARG, "from_s3tc: buffers must be %zu-byte aligned", align);
	if (int r = ensure_device()) return r;
	const size_t groups = (nblocks + 255) / 256;
	if (groups > 0xffffffffull) return set_error(S2TC_E_RANGE, "from_s3tc: too many blocks");
	hipStream_t st = (hipStream_t)stream;
	if (dxt == 0) k_from_s3tc<0><<<(unsigned)groups, 256, 0, st>>>(d_in, d_out, nblocks);
	else if (dxt == 1) k_from_s3tc<1><<<(unsigned)groups, 256, 0, st>>>(d_in, d_out, nblocks);
	else k_from_s3tc<2><<<(unsigned)groups, 256, 0, st>>>(d_in, d_out, nblocks);
	return hip_check(hipGetLastError(), "launch k_from_s3tc");
}
