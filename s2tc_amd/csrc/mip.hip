// mip.hip - the CLI's mip chain on the device.
//
// Image_MipReduce32 (s2tc_compress.c:427-493) halves an RGBA8 image with a
// truncating box filter: both sides > 1 -> (a + b + c + d) >> 2 over 2x2,
// only the width > 1 -> (a + b) >> 1 over horizontal pairs, only the height
// > 1 -> (a + b) >> 1 over vertical pairs; odd last rows / columns are
// dropped (the new size is the old size >> 1).  The chain of
// s2tc_compress.c:722-733 repeats it down to 1x1.  One thread per output
// texel; the four channels are summed in 16-bit lanes of two words.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "s2tc_internal.h"

namespace s2tc {

// per-channel sum of RGBA8 words in two words of 16-bit lanes (r|b, g|a)
struct Sum4 {
	uint32_t rb, ga;
	__device__ __forceinline__ void add(uint32_t p)
	{
		rb += p & 0x00ff00ffu;
		ga += (p >> 8) & 0x00ff00ffu;
	}
	__device__ __forceinline__ uint32_t shifted(int s) const
	{
		return ((rb >> s) & 0x00ff00ffu) | (((ga >> s) & 0x00ff00ffu) << 8);
	}
};

__device__ __forceinline__ void mip_texel(const uint32_t *__restrict__ in, int width, int height,
                                          uint32_t *__restrict__ out, int x, int y, int ow);

__global__ __launch_bounds__(256) void k_mip_reduce(const uint32_t *__restrict__ in, int width, int height,
                                                    uint32_t *__restrict__ out)
{
	const int ow = width > 1 ? width >> 1 : 1, oh = height > 1 ? height >> 1 : 1;
	const int x = (int)(blockIdx.x * 64 + threadIdx.x), y = (int)(blockIdx.y * 4 + threadIdx.y);
	if (x >= ow || y >= oh) return;
	mip_texel(in, width, height, out, x, y, ow);
}

// One level of many chains in one launch: a task is one texture's halving
// at this level; tasks own consecutive runs of 256-texel tiles (tile0 is the
// first), and a workgroup finds its task by a binary search over tile0.
struct MipTask {
	const uint32_t *in;
	uint32_t *out;
	int w, h;          // input size
	uint32_t tile0;
	uint32_t pad;
};

__device__ __forceinline__ void mip_texel(const uint32_t *__restrict__ in, int width, int height,
                                          uint32_t *__restrict__ out, int x, int y, int ow)
{
	Sum4 s{0u, 0u};
	int shift;
	if (width > 1 && height > 1) {
		const uint32_t *r0 = in + (size_t)(2 * y) * (size_t)width + 2 * x, *r1 = r0 + width;
		s.add(r0[0]);
		s.add(r0[1]);
		s.add(r1[0]);
		s.add(r1[1]);
		shift = 2;
	} else if (width > 1) {
		const uint32_t *r0 = in + (size_t)y * (size_t)width + 2 * x;
		s.add(r0[0]);
		s.add(r0[1]);
		shift = 1;
	} else {
		const uint32_t *r0 = in + (size_t)(2 * y) * (size_t)width + x;
		s.add(r0[0]);
		s.add(r0[width]);
		shift = 1;
	}
	out[(size_t)y * (size_t)ow + x] = s.shifted(shift);
}

__global__ __launch_bounds__(256) void k_mip_reduce_batch(const MipTask *__restrict__ tasks, int ntasks)
{
	const uint32_t tile = blockIdx.x;
	int lo = 0, hi = ntasks - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (tasks[mid].tile0 <= tile) lo = mid;
		else hi = mid - 1;
	}
	const MipTask t = tasks[lo];
	const int ow = t.w > 1 ? t.w >> 1 : 1, oh = t.h > 1 ? t.h >> 1 : 1;
	const uint32_t i = (tile - t.tile0) * 256u + threadIdx.x;
	if (i >= (uint32_t)ow * (uint32_t)oh) return;
	const int y = (int)(i / (uint32_t)ow), x = (int)(i - (uint32_t)y * (uint32_t)ow);
	mip_texel(t.in, t.w, t.h, t.out, x, y, ow);
}

}  // namespace s2tc

using namespace s2tc;

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

extern "C" int s2tc_amd_mip_layout(int width, int height, int *nlevels, size_t *offsets, int *widths, int *heights,
                                   int max_levels, size_t *total_bytes)
{
	if (width <= 0 || height <= 0 || !nlevels) return set_error(S2TC_E_ARG, "mip_layout: bad arguments");
	int w = width, h = height, n = 0;
	size_t off = 0;
	for (;;) {
		if (n < max_levels) {
			if (offsets) offsets[n] = off;
			if (widths) widths[n] = w;
			if (heights) heights[n] = h;
		}
		off += align256((size_t)w * h * 4);
		++n;
		if (w == 1 && h == 1) break;
		w = w > 1 ? w >> 1 : 1;
		h = h > 1 ? h >> 1 : 1;
	}
	*nlevels = n;
	if (total_bytes) *total_bytes = off;
	return 0;
}

extern "C" int s2tc_amd_mip_reduce(const void *d_in, int width, int height, void *d_out, void *stream)
{
	if (!d_in || !d_out || width <= 0 || height <= 0) return set_error(S2TC_E_ARG, "mip_reduce: bad arguments");
	if (width == 1 && height == 1) return set_error(S2TC_E_ARG, "mip_reduce: the image is already 1x1");
	if (int r = ensure_device()) return r;
	const int ow = width > 1 ? width >> 1 : 1, oh = height > 1 ? height >> 1 : 1;
	if ((oh + 3) / 4 > 65535) return set_error(S2TC_E_RANGE, "mip_reduce: image too tall");
	dim3 grid((unsigned)((ow + 63) / 64), (unsigned)((oh + 3) / 4)), block(64, 4);
	k_mip_reduce<<<grid, block, 0, (hipStream_t)stream>>>((const uint32_t *)d_in, width, height, (uint32_t *)d_out);
	return hip_check(hipGetLastError(), "launch k_mip_reduce");
}

extern "C" int s2tc_amd_mip_chain(const void *d_src, int width, int height, void *d_levels, size_t capacity,
                                  void *stream)
{
	int n = 0;
	size_t total = 0;
	if (int r = s2tc_amd_mip_layout(width, height, &n, nullptr, nullptr, nullptr, 0, &total)) return r;
	if (!d_src || !d_levels) return set_error(S2TC_E_ARG, "mip_chain: null buffer");
	if (capacity < total) return set_error(S2TC_E_ARG, "mip_chain: %zu bytes needed, %zu given", total, capacity);
	if (int r = ensure_device()) return r;
	hipStream_t st = (hipStream_t)stream;
	uint8_t *base = (uint8_t *)d_levels;
	int rc = 0;
	if (d_src != d_levels)
		rc = hip_check(hipMemcpyAsync(base, d_src, (size_t)width * height * 4, hipMemcpyDeviceToDevice, st), "D2D");
	int w = width, h = height;
	size_t off = 0;
	while (!rc && !(w == 1 && h == 1)) {
		const size_t next = off + align256((size_t)w * h * 4);
		rc = s2tc_amd_mip_reduce(base + off, w, h, base + next, st);
		off = next;
		w = w > 1 ? w >> 1 : 1;
		h = h > 1 ? h >> 1 : 1;
	}
	return rc ? rc : n;
}

extern "C" int s2tc_amd_mip_chain_batch(const s2tc_amd_mip_job *jobs, int njobs, void *stream)
{
	if (njobs < 0 || (njobs && !jobs)) return set_error(S2TC_E_ARG, "mip_chain_batch: bad arguments");
	// per texture: level offsets and sizes; then the tasks of every level
	std::vector<std::vector<MipTask>> levels;
	for (int j = 0; j < njobs; ++j) {
		const s2tc_amd_mip_job &jb = jobs[j];
		if (!jb.levels || jb.width <= 0 || jb.height <= 0)
			return set_error(S2TC_E_ARG, "mip_chain_batch: job %d: bad image", j);
		int w = jb.width, h = jb.height, l = 0;
		size_t off = 0;
		while (!(w == 1 && h == 1)) {
			const size_t next = off + align256((size_t)w * h * 4);
			if ((size_t)l >= levels.size()) levels.emplace_back();
			MipTask t;
			t.in = (const uint32_t *)((const uint8_t *)jb.levels + off);
			t.out = (uint32_t *)((uint8_t *)jb.levels + next);
			t.w = w;
			t.h = h;
			t.tile0 = 0;
			t.pad = 0;
			levels[l].push_back(t);
			off = next;
			w = w > 1 ? w >> 1 : 1;
			h = h > 1 ? h >> 1 : 1;
			++l;
		}
	}
	if (levels.empty()) return 0;
	std::vector<size_t> first(levels.size());
	std::vector<uint32_t> ntiles(levels.size());
	std::vector<MipTask> all;
	for (size_t l = 0; l < levels.size(); ++l) {
		uint64_t tiles = 0;
		for (MipTask &t : levels[l]) {
			t.tile0 = (uint32_t)tiles;
			const uint64_t ow = t.w > 1 ? t.w >> 1 : 1, oh = t.h > 1 ? t.h >> 1 : 1;
			tiles += (ow * oh + 255) / 256;
		}
		if (tiles > 0x7fffffffull) return set_error(S2TC_E_RANGE, "mip_chain_batch: level %zu too large", l);
		first[l] = all.size();
		ntiles[l] = (uint32_t)tiles;
		all.insert(all.end(), levels[l].begin(), levels[l].end());
	}
	if (int r = ensure_device()) return r;
	hipStream_t st = (hipStream_t)stream;
	MipTask *d = nullptr;
	int rc = hip_check(hipMalloc((void **)&d, all.size() * sizeof(MipTask)), "hipMalloc mip tasks");
	if (!rc)
		rc = hip_check(hipMemcpyAsync(d, all.data(), all.size() * sizeof(MipTask), hipMemcpyHostToDevice, st),
		               "H2D mip tasks");
	for (size_t l = 0; !rc && l < levels.size(); ++l) {
		k_mip_reduce_batch<<<ntiles[l], 256, 0, st>>>(d + first[l], (int)levels[l].size());
		rc = hip_check(hipGetLastError(), "launch k_mip_reduce_batch");
	}
	// the task list is read by the kernels: keep it until they are done
	const int rs = hip_check(hipStreamSynchronize(st), "mip_chain_batch synchronize");
	if (d) (void)hipFree(d);
	return rc ? rc : rs;
}
