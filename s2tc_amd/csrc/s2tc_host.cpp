// s2tc_host.cpp - host side of the GPU encoder: kernel selection, launch
// geometry, device memory plumbing and error reporting.
//
// The dispatch rules mirror s2tc_encode_block_func (s2tc_algorithm.cpp
// :1110-1194): out-of-range refine -> ALWAYS, dxt -> DXT5, distance -> WAVG;
// quick mode (one lane per block) iff nrandom < 0 and the distance is not
// NORMALMAP; MODE_NORMAL without random colours (nrandom <= 0) runs the
// one-lane-per-block all-16 kernel, MODE_NORMAL with random colours the
// 16-lanes-per-block kernel.
#include "plan.h"
#include "s2tc_internal.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>

namespace s2tc {

static thread_local char g_err[512];

int set_error(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof g_err, fmt, ap);
	va_end(ap);
	return code;
}

int hip_check(hipError_t e, const char *what)
{
	if (e == hipSuccess) return 0;
	return set_error(S2TC_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

static KernelTable g_tables[8];
static std::once_flag g_tables_once;

const KernelTable &kernel_table(int cd)
{
	std::call_once(g_tables_once, [] {
		s2tc_kernels_cd0(&g_tables[0]);
		s2tc_kernels_cd1(&g_tables[1]);
		s2tc_kernels_cd2(&g_tables[2]);
		s2tc_kernels_cd3(&g_tables[3]);
		s2tc_kernels_cd4(&g_tables[4]);
		s2tc_kernels_cd5(&g_tables[5]);
		s2tc_kernels_cd6(&g_tables[6]);
		s2tc_kernels_cd7(&g_tables[7]);
	});
	return g_tables[cd];
}

void normalize_modes(int &dxt, int &cd, int &refine)
{
	if (refine != REFINE_NEVER && refine != REFINE_LOOP) refine = REFINE_ALWAYS;
	if (dxt != DXT1 && dxt != DXT3) dxt = DXT5;
	if (cd < CD_RGB || cd > CD_NORMALMAP) cd = CD_WAVG;
}

bool is_fast(int cd, int nrandom) { return cd != CD_NORMALMAP && nrandom < 0; }

int normal_pbw(int nrandom) { return (16 + (nrandom > 0 ? nrandom : 0) + 15) / 16; }

// Per-device tables (the NORMALMAP feature table) are built on the first
// call that reaches a device, outside any stream capture: callers that
// capture graphs make one ordinary call first.
static std::mutex g_init_mu;
static std::atomic<uint64_t> g_init_done{0};

static int ensure_tables()
{
	int dev = 0;
	if (int r = hip_check(hipGetDevice(&dev), "hipGetDevice")) return r;
	const uint64_t bit = dev < 64 ? 1ull << dev : 0ull;
	if (bit && (g_init_done.load(std::memory_order_acquire) & bit)) return 0;
	std::lock_guard<std::mutex> lk(g_init_mu);
	if (bit && (g_init_done.load(std::memory_order_acquire) & bit)) return 0;
	hipStream_t st = nullptr;
	if (int r = hip_check(hipStreamCreate(&st), "hipStreamCreate")) return r;
	int rc = 0;
	for (int cd = 0; cd < 8 && !rc; ++cd) {
		const KernelTable &kt = kernel_table(cd);
		if (!kt.init) continue;
		int dummy = 0;
		void *args[] = {&dummy};
		rc = hip_check(hipLaunchKernel(kt.init, dim3(256), dim3(256), args, 0, st), "launch table set-up");
	}
	if (!rc) rc = hip_check(hipStreamSynchronize(st), "table set-up");
	(void)hipStreamDestroy(st);
	if (!rc && bit) g_init_done.fetch_or(bit, std::memory_order_acq_rel);
	return rc;
}

int ensure_device()
{
	static int ok = -1;
	if (ok < 0) {
		int n = 0;
		hipError_t e = hipGetDeviceCount(&n);
		ok = (e == hipSuccess && n > 0) ? 1 : 0;
	}
	if (!ok) return set_error(S2TC_E_NOGPU, "no HIP device available");
	return ensure_tables();
}

int launch_encode(const EncodeArgs &a, hipStream_t stream, void *ext_scratch, size_t ext_bytes)
{
	int dxt = a.dxt, cd = a.cd, refine = a.refine;
	normalize_modes(dxt, cd, refine);
	if (a.width <= 0 || a.height <= 0) return set_error(S2TC_E_ARG, "bad image size %dx%d", a.width, a.height);
	if (a.src_kind != S2TC_SRC_QUANTIZED && a.src_kind != S2TC_SRC_RGBA8)
		return set_error(S2TC_E_ARG, "bad source kind %d", a.src_kind);
	if (a.nrandom > S2TC_MAX_NRANDOM)
		return set_error(S2TC_E_RANGE, "nrandom %d exceeds the GPU limit %d", a.nrandom, S2TC_MAX_NRANDOM);
	if (int r = ensure_device()) return r;
	const KernelTable &kt = kernel_table(cd);
	const int raw = a.src_kind == S2TC_SRC_RGBA8 ? 1 : 0;
	const long long nb = (long long)((a.width + 3) / 4) * ((a.height + 3) / 4);
	if (nb > (1LL << 30)) return set_error(S2TC_E_RANGE, "image too large");
	const void *src = a.src;
	void *dst = a.dst;
	int width = a.width, height = a.height;
	if (is_fast(cd, a.nrandom) || a.nrandom <= 0) {
		// one lane per block: quick mode, or only the block's colours
		const void *k = is_fast(cd, a.nrandom) ? kt.fast[dxt][raw] : kt.all16[dxt][raw];
		if (!k) return set_error(S2TC_E_ARG, "no one-lane kernel for distance %d", cd);
		int vec = ((width & 3) == 0 && (((uintptr_t)src) & 15) == 0) ? 1 : 0;
		const int bw = (width + 3) / 4, bh = (height + 3) / 4;
		const int rows_per_launch = 4 * 65535;  // grid.y limit of 65535 workgroups of 4 block rows
		int first_general_row = 0;
		const void *ki = is_fast(cd, a.nrandom) ? kt.fast_int[dxt][raw][refine] : nullptr;
		if (vec && ki) {
			// full blocks of 16-byte rows first, then the partial bottom block row
			int nbx = width / 4, nby = height / 4;
			for (int by0 = 0; by0 < nby; by0 += rows_per_launch) {
				const int rows = nby - by0 < rows_per_launch ? nby - by0 : rows_per_launch;
				void *args[] = {(void *)&src, (void *)&width, (void *)&nbx, (void *)&nby, (void *)&dst, (void *)&by0};
				dim3 grid((unsigned)((nbx + 63) / 64), (unsigned)((rows + 3) / 4)), block(64, 4);
				if (int r = hip_check(hipLaunchKernel(ki, grid, block, args, 0, stream), "launch k_fast_interior"))
					return r;
			}
			first_general_row = nby;
		}
		for (int by0 = first_general_row; by0 < bh; by0 += rows_per_launch) {
			const int rows = bh - by0 < rows_per_launch ? bh - by0 : rows_per_launch;
			void *args[] = {(void *)&src, (void *)&width, (void *)&height, (void *)&dst,
			                (void *)&refine, (void *)&vec, (void *)&by0};
			dim3 grid((unsigned)((bw + 63) / 64), (unsigned)((rows + 3) / 4)), block(64, 4);
			if (int r = hip_check(hipLaunchKernel(k, grid, block, args, 0, stream), "launch one-lane encoder")) return r;
		}
		return 0;
	}
	int nrandom = a.nrandom;
	int pbw = normal_pbw(nrandom);
	unsigned long long seed = a.seed;
	unsigned block_base = a.block_index_base;
	if (nrandom <= S2TC_LDS_NRANDOM) {
		size_t smem = normal_group_smem(kt, pbw);
		void *args[] = {(void *)&src, (void *)&width, (void *)&height, (void *)&dst, (void *)&dxt,
		                (void *)&nrandom, (void *)&refine, (void *)&seed, (void *)&pbw, (void *)&block_base};
		dim3 grid((unsigned)((nb + 3) / 4)), block(64);
		return hip_check(hipLaunchKernel(kt.normal[raw], grid, block, args, smem, stream), "launch k_normal");
	}
	// more candidates than LDS holds: group memory in the caller's scratch, or
	// in stream-ordered scratch that lives exactly as long as this launch
	int grid = 0;
	size_t bytes = 0;
	wide_launch_shape(kt, pbw, (nb + 3) / 4, device_cu_count(), grid, bytes);
	void *gmem = ext_scratch;
	if (ext_scratch) {
		const size_t wg = bytes / (size_t)grid;
		if (ext_bytes < wg)
			return set_error(S2TC_E_RANGE, "nrandom %d needs %zu bytes of scratch per workgroup, %zu available",
			                 nrandom, wg, ext_bytes);
		grid = (int)std::min((size_t)grid, ext_bytes / wg);
	} else if (hipMallocAsync(&gmem, bytes, stream) != hipSuccess) {
		return set_error(S2TC_E_RANGE, "cannot allocate %zu bytes of scratch for nrandom %d", bytes, nrandom);
	}
	void *args[] = {(void *)&src, (void *)&width, (void *)&height, (void *)&dst, (void *)&dxt, (void *)&nrandom,
	                (void *)&refine, (void *)&seed, (void *)&pbw, (void *)&block_base, (void *)&gmem};
	const int rc = hip_check(hipLaunchKernel(kt.normal_wide[raw], dim3((unsigned)grid), dim3(64), args, 0, stream),
	                         "launch k_normal_wide");
	if (ext_scratch) return rc;
	const int rc2 = hip_check(hipFreeAsync(gmem, stream), "hipFreeAsync (scratch)");
	return rc ? rc : rc2;
}

}  // namespace s2tc

using namespace s2tc;

extern "C" {

const char *s2tc_amd_last_error(void) { return g_err; }

int s2tc_amd_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) return 0;
	return n;
}

int s2tc_amd_malloc(void **dptr, size_t bytes)
{
	if (!dptr) return set_error(S2TC_E_ARG, "null pointer");
	if (int r = ensure_device()) return r;
	return hip_check(hipMalloc(dptr, bytes ? bytes : 1), "hipMalloc");
}

int s2tc_amd_free(void *dptr) { return hip_check(hipFree(dptr), "hipFree"); }

int s2tc_amd_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
	return hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

int s2tc_amd_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
	return hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

int s2tc_amd_memset(void *dst, int value, size_t bytes)
{
	return hip_check(hipMemset(dst, value, bytes), "hipMemset");
}

int s2tc_amd_synchronize(void) { return hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int s2tc_amd_encode_image(const void *d_src, int src_kind, int width, int height, void *d_dst, int dxt, int cd,
                          int nrandom, int refine, uint64_t seed, void *stream)
{
	if (!d_src || !d_dst) return set_error(S2TC_E_ARG, "null device pointer");
	EncodeArgs a{d_src, d_dst, src_kind, width, height, dxt, cd, nrandom, refine, seed, 0};
	return launch_encode(a, (hipStream_t)stream);
}

int s2tc_amd_debug_set_scan_cap(unsigned cap)
{
	if (int r = ensure_device()) return r;
	for (int cd = 0; cd < 8; ++cd) {
		const KernelTable &kt = kernel_table(cd);
		if (!kt.set_scan_cap) continue;
		void *args[] = {&cap};
		if (int r = hip_check(hipLaunchKernel(kt.set_scan_cap, dim3(1), dim3(64), args, 0, nullptr),
		                      "launch k_set_scan_cap"))
			return r;
	}
	return hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

}  // extern "C"
