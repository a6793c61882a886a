// util.hip - a plain streaming copy kernel (for calibrating byte counters:
// it reads and writes exactly `bytes` bytes once each).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "s2tc_internal.h"

namespace s2tc {

__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
		dst[i] = src[i];
}

}  // namespace s2tc

extern "C" int s2tc_amd_memcpy_d2d_kernel(void *dst, const void *src, size_t bytes, void *stream)
{
	using namespace s2tc;
	if (!dst || !src || (bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15))
		return set_error(S2TC_E_ARG, "copy kernel needs 16-byte aligned pointers and sizes");
	const size_t n16 = bytes / 16;
	const unsigned grid = (unsigned)((n16 + 255) / 256 < 65536 ? (n16 + 255) / 256 : 65536);
	k_copy16<<<grid ? grid : 1, 256, 0, (hipStream_t)stream>>>((const uint4 *)src, (uint4 *)dst, n16);
	return hip_check(hipGetLastError(), "copy kernel");
}
