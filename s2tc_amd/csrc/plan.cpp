// plan.cpp - grouping of images into kernel classes (see plan.h).
#include "plan.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace s2tc {

int device_cu_count()
{
	static int ncu = 0;
	if (!ncu) {
		int dev = 0;
		(void)hipGetDevice(&dev);
		hipDeviceProp_t prop;
		ncu = (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
		          ? prop.multiProcessorCount
		          : 256;
	}
	return ncu;
}

// tile -> (first job, last job) of tiles of `tile` packed blocks, by a walk
// over the jobs' first blocks
static void tile_ranges(const std::vector<JobDesc> &jobs, long long nblocks, int tile, int ntiles,
                        std::vector<TileJobs> &out)
{
	out.resize((size_t)ntiles);
	int jlo = 0;
	const int nj = (int)jobs.size();
	for (int t = 0; t < ntiles; ++t) {
		const long long b0 = (long long)t * tile, b1 = std::min(nblocks, b0 + tile) - 1;
		while (jlo + 1 < nj && jobs[jlo + 1].block0 <= b0) ++jlo;
		int jhi = jlo;
		while (jhi + 1 < nj && jobs[jhi + 1].block0 <= b1) ++jhi;
		out[(size_t)t] = TileJobs{jlo, jhi};
	}
}

int plan_classes(const s2tc_amd_job *jobs, int njobs, std::vector<ClassPlan> &classes, bool with_forms)
{
	classes.clear();
	// class slots: 0..71 quick ((cd*3 + dxt)*3 + refine), 72..79 MODE_NORMAL
	// with random colours (cd), 80..151 MODE_NORMAL without
	// (80 + (cd*3 + dxt)*3 + refine), 152..159 MODE_NORMAL with more random
	// colours than LDS holds (152 + cd)
	int slot[160];
	for (int &v : slot) v = -1;
	for (int i = 0; i < njobs; ++i) {
		const s2tc_amd_job &j = jobs[i];
		int dxt = j.dxt, cd = j.cd, refine = j.refine;
		normalize_modes(dxt, cd, refine);
		if (j.width <= 0 || j.height <= 0 || !j.src || !j.dst) return set_error(S2TC_E_ARG, "job %d: bad image", i);
		// block indices are int and the kernels split them with 24-bit multiplies
		const long long jbw = (j.width + 3LL) / 4, jbh = (j.height + 3LL) / 4;
		if (jbw >= (1LL << 23) || jbh >= (1LL << 23) || jbw * jbh >= (1LL << 31))
			return set_error(S2TC_E_RANGE, "job %d: %d x %d texels exceed the batch limits", i, j.width, j.height);
		if (j.nrandom > S2TC_MAX_NRANDOM)
			return set_error(S2TC_E_RANGE, "job %d: nrandom %d exceeds the GPU limit %d", i, j.nrandom,
			                 S2TC_MAX_NRANDOM);
		if (j.src_kind != S2TC_SRC_QUANTIZED && j.src_kind != S2TC_SRC_RGBA8)
			return set_error(S2TC_E_ARG, "job %d: bad source kind", i);
		const bool fast = is_fast(cd, j.nrandom);
		const bool all16 = !fast && j.nrandom <= 0;
		const bool wide = !fast && j.nrandom > S2TC_LDS_NRANDOM;
		const int key = fast ? (cd * 3 + dxt) * 3 + refine
		                : all16 ? 80 + (cd * 3 + dxt) * 3 + refine
		                : wide ? 152 + cd : 72 + cd;
		if (slot[key] < 0) {
			slot[key] = (int)classes.size();
			classes.emplace_back();
			ClassPlan &c = classes.back();
			c.normal = !fast && !all16;
			c.wide = wide;
			c.all16 = all16;
			c.cd = cd;
			c.dxt = c.normal ? -1 : dxt;
			c.refine = (fast || all16) ? refine : -1;
			c.tile = c.normal ? 4 : 256;
		}
		ClassPlan &c = classes[slot[key]];
		JobDesc d;
		d.src = (const uint32_t *)j.src;
		d.dst = (uint8_t *)j.dst;
		d.width = j.width;
		d.height = j.height;
		d.dxt = dxt;
		d.nrandom = fast ? -1 : (j.nrandom > 0 ? j.nrandom : 0);
		d.refine = refine;
		d.raw = j.src_kind == S2TC_SRC_RGBA8 ? 1 : 0;
		d.seed = j.seed;
		d.block0 = c.nblocks;
		d.bw_magic = bw_magic((j.width + 3) / 4);
		d.pad = 0;
		const long long nb = (long long)((j.width + 3) / 4) * ((j.height + 3) / 4);
		c.nblocks += nb;
		c.texels += (long long)j.width * j.height;
		c.bytes += (long long)j.width * j.height * 4 + nb * block_bytes(dxt);
		c.pbw_max = std::max(c.pbw_max, normal_pbw(d.nrandom));
		c.jobs.push_back(d);
		c.job_index.push_back(i);
		// rough cost: quick mode ~1 per block, MODE_NORMAL ~ candidate pairs
		const int m = 16 + std::max(0, d.nrandom);
		c.cost += fast ? (double)nb : all16 ? (double)nb * 8.0 * (dxt == DXT5 ? 2 : 1)
		                                    : (double)nb * (m * (m - 1) / 2) * (dxt == DXT5 ? 2 : 1) / 8.0;
	}
	for (auto &c : classes) {
		const KernelTable &kt = kernel_table(c.cd);
		c.kernel = c.wide     ? kt.normal_batch_wide
		           : c.normal ? kt.normal_batch
		           : c.all16  ? kt.all16_batch[c.dxt][c.refine]
		                      : kt.fast_batch[c.dxt][c.refine];
		if (!c.kernel) return set_error(S2TC_E_ARG, "no batch kernel for distance %d", c.cd);
		c.ntiles = (int)((c.nblocks + c.tile - 1) / c.tile);
		c.smem = c.normal && !c.wide ? normal_group_smem(kt, c.pbw_max) : 0;
		if (c.wide) wide_launch_shape(kt, c.pbw_max, c.ntiles, device_cu_count(), c.wide_grid, c.scratch);
		tile_ranges(c.jobs, c.nblocks, c.tile, c.ntiles, c.tiles);
		if (c.all16) {
			c.forms[FORM_LANE] = c.kernel;
			const void *k2 = kt.all16_batch_alone[c.dxt][c.refine];
			c.forms[FORM_LANE_2WAVES] = k2 != c.kernel ? k2 : nullptr;
			if (with_forms) {
				// the random-candidate kernel runs nrandom-0 jobs as they are
				c.forms[FORM_ROW16] = kt.normal_batch;
				c.ntiles4 = (int)((c.nblocks + 3) / 4);
				c.smem4 = normal_group_smem(kt, 1);
				tile_ranges(c.jobs, c.nblocks, 4, c.ntiles4, c.tiles4);
			}
		}
	}
	std::stable_sort(classes.begin(), classes.end(),
	                 [](const ClassPlan &x, const ClassPlan &y) { return x.cost > y.cost; });
	return 0;
}

static size_t align16(size_t n) { return (n + 15) & ~(size_t)15; }

size_t plan_descriptor_bytes(const std::vector<ClassPlan> &classes)
{
	size_t n = 0;
	for (const auto &c : classes)
		n += align16(c.jobs.size() * sizeof(JobDesc)) + align16(c.tiles.size() * sizeof(TileJobs)) +
		     align16(c.tiles4.size() * sizeof(TileJobs));
	return n;
}

void plan_pack(std::vector<ClassPlan> &classes, uint8_t *host, uint8_t *dev)
{
	size_t off = 0;
	for (auto &c : classes) {
		const size_t jb = c.jobs.size() * sizeof(JobDesc), tb = c.tiles.size() * sizeof(TileJobs);
		memcpy(host + off, c.jobs.data(), jb);
		c.d_jobs = (JobDesc *)(dev + off);
		off += align16(jb);
		memcpy(host + off, c.tiles.data(), tb);
		c.d_tiles = (TileJobs *)(dev + off);
		off += align16(tb);
		const size_t t4 = c.tiles4.size() * sizeof(TileJobs);
		if (t4) memcpy(host + off, c.tiles4.data(), t4);
		c.d_tiles4 = t4 ? (TileJobs *)(dev + off) : nullptr;
		off += align16(t4);
	}
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

size_t plan_scratch_bytes(const std::vector<ClassPlan> &classes)
{
	size_t n = 0;
	for (const auto &c : classes) n += align256(c.scratch);
	return n;
}

void plan_set_scratch(std::vector<ClassPlan> &classes, uint8_t *dev)
{
	size_t off = 0;
	for (auto &c : classes) {
		c.d_scratch = c.scratch ? dev + off : nullptr;
		off += align256(c.scratch);
	}
}

void wide_launch_shape(const KernelTable &kt, int pbw, long long ntiles, int ncu, int &grid, size_t &bytes)
{
	const size_t wg = normal_group_smem(kt, pbw);
	const long long cap = std::max(1LL, (long long)(((size_t)1 << 30) / wg));
	grid = (int)std::max(1LL, std::min(ntiles, std::min(4LL * ncu, cap)));
	bytes = wg * (size_t)grid;
}

// Workgroups of `kernel` that one CU holds at once with `smem` bytes of
// dynamic LDS each (registers, LDS and wave slots), cached per shape.
static int resident_per_cu(const void *kernel, int block, size_t smem)
{
	struct Entry {
		const void *k;
		int block;
		size_t smem;
		int n;
	};
	static std::mutex mu;
	static std::vector<Entry> cache;
	std::lock_guard<std::mutex> lock(mu);
	for (const Entry &e : cache)
		if (e.k == kernel && e.block == block && e.smem == smem) return e.n;
	int n = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, smem) != hipSuccess || n < 1) n = 0;
	cache.push_back(Entry{kernel, block, smem, n});
	return n;
}

// Grid of a persistent launch: `want` workgroups rounded down to whole rounds
// of what the device holds at once.  The kernels stride over their tiles by
// the grid size, so each workgroup's share is fixed: a partial last round
// (8 wanted where 7 fit a CU) leaves its workgroups to run whole shares after
// the others, stretching the launch, while whole rounds (8 wanted where 2
// fit) let the dispatcher rebalance the CUs between rounds.
static int persistent_grid(const void *kernel, int block, size_t smem, long long want, int ncu)
{
	const int r = resident_per_cu(kernel, block, smem);
	const long long cap = (long long)r * ncu;
	if (r > 0 && want > cap) want = (want / cap) * cap;
	return (int)std::max(1LL, want);
}

int plan_launch_class(const ClassPlan &c, int wg_per_cu, int ncu, hipStream_t st)
{
	long long nb = c.nblocks;
	int ntiles = c.ntiles;
	int grid = wg_per_cu <= 0 ? ntiles : std::min(ntiles, wg_per_cu * ncu);
	if (grid < 1) return 0;
	const JobDesc *jobs = c.d_jobs;
	const TileJobs *tiles = c.d_tiles;
	if (c.wide) {
		int pbw = c.pbw_max;
		uint8_t *gmem = c.d_scratch;
		if (!gmem) return set_error(S2TC_E_ARG, "wide class without scratch memory");
		void *args[] = {(void *)&jobs, (void *)&tiles, (void *)&nb, (void *)&ntiles, (void *)&pbw, (void *)&gmem};
		return hip_check(hipLaunchKernel(c.kernel, dim3(c.wide_grid), dim3(64), args, 0, st),
		                 "launch k_normal_batch_wide");
	}
	if (c.all16 && c.form == FORM_ROW16) {
		int pbw = 1, nt4 = c.ntiles4;
		const TileJobs *t4 = c.d_tiles4;
		if (!t4 || !c.forms[FORM_ROW16]) return set_error(S2TC_E_ARG, "class has no 16-lane form");
		const int grid4 = wg_per_cu <= 0 ? nt4
		                                 : persistent_grid(c.forms[FORM_ROW16], 64, c.smem4,
		                                                   std::min(nt4, 4 * wg_per_cu * ncu), ncu);
		void *args[] = {(void *)&jobs, (void *)&t4, (void *)&nb, (void *)&nt4, (void *)&pbw};
		return hip_check(hipLaunchKernel(c.forms[FORM_ROW16], dim3(grid4), dim3(64), args, c.smem4, st),
		                 "launch k_normal_batch (all-16 jobs)");
	}
	if (c.normal) {
		int pbw = c.pbw_max;
		if (wg_per_cu > 0) grid = persistent_grid(c.kernel, 64, c.smem, grid, ncu);
		void *args[] = {(void *)&jobs, (void *)&tiles, (void *)&nb, (void *)&ntiles, (void *)&pbw};
		return hip_check(hipLaunchKernel(c.kernel, dim3(grid), dim3(64), args, c.smem, st), "launch k_normal_batch");
	}
	void *args[] = {(void *)&jobs, (void *)&tiles, (void *)&nb, (void *)&ntiles};
	const void *k = c.all16 && c.forms[c.form] ? c.forms[c.form] : c.kernel;
	if (wg_per_cu > 0) grid = persistent_grid(k, 256, 0, grid, ncu);
	return hip_check(hipLaunchKernel(k, dim3(grid), dim3(256), args, 0, st),
	                 c.all16 ? "launch k_all16_batch" : "launch k_fast_batch");
}


}  // namespace s2tc
