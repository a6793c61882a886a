// noalloc.hip - s2tc_amd_encode_image_noalloc: s2tc_amd_encode_image without
// any device memory allocation, for callers that cannot allocate (code under
// stream capture, tight serving loops, hosts that forbid hipMalloc).
//
// The only kernels of the encode that need memory beyond the source and the
// blocks are the random-candidate kernels above S2TC_LDS_NRANDOM (candidate
// lists, distance rows from the 33rd on and column features of four blocks
// per workgroup, under 4 * m * 68 bytes).  Here that group memory comes from
// `scratch`, a static device array of S2TC_NOALLOC_SCRATCH bytes that every
// GPU holds from module load to process end; the persistent grid shrinks to
// the workgroups whose slices fit in it.  Nothing in the array needs to start
// at zero: each workgroup writes its candidates and rows before reading them.
// Calls that overlap on one GPU (concurrent streams or host threads) share
// the array and race; calls on one stream, or on different GPUs, do not.
#include "s2tc_internal.h"

#include <hip/hip_runtime.h>

#include <mutex>

namespace {

static __device__ __align__(256) uint8_t scratch[S2TC_NOALLOC_SCRATCH];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch` on each device, set on first use

// The device address of this device's copy of `scratch`.  The array's name is
// only a host-side handle: kernels and copies must use this address.
int scratch_address(void **out)
{
	int dev = 0;
	if (int r = s2tc::hip_check(hipGetDevice(&dev), "hipGetDevice")) return r;
	if (dev < 0 || dev >= kMaxDevices) return s2tc::set_error(S2TC_E_RANGE, "device %d beyond %d", dev, kMaxDevices);
	std::lock_guard<std::mutex> lk(g_mu);
	if (!g_ptr[dev]) {
		void *p = nullptr;
		if (int r = s2tc::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress (scratch)"))
			return r;
		if (!p) return s2tc::set_error(S2TC_E_HIP, "no device address for the scratch array");
		g_ptr[dev] = p;
	}
	*out = g_ptr[dev];
	return 0;
}

}  // namespace

extern "C" int s2tc_amd_encode_image_noalloc(const void *d_src, int src_kind, int width, int height, void *d_dst,
                                             int dxt, int cd, int nrandom, int refine, uint64_t seed, void *stream)
{
	if (!d_src || !d_dst) return s2tc::set_error(S2TC_E_ARG, "null device pointer");
	if (int r = s2tc::ensure_device()) return r;
	void *ptr = nullptr;
	if (int r = scratch_address(&ptr)) return r;
	s2tc::EncodeArgs a{d_src, d_dst, src_kind, width, height, dxt, cd, nrandom, refine, seed, 0};
	return s2tc::launch_encode(a, (hipStream_t)stream, ptr, S2TC_NOALLOC_SCRATCH);
}
