// batch.cpp - many images (textures x mip levels) in one hipGraph launch.
//
// Plan: jobs are grouped into kernel classes.  Quick-mode jobs (nrandom < 0,
// distance != NORMALMAP) form one class per (distance, format) and run
// k_fast_batch (one lane per 4x4 block, tiles of 256 blocks); MODE_NORMAL
// jobs without random colours form one class per (distance, format) and run
// k_all16_batch (also one lane per block); jobs with random colours form one
// class per distance and run k_normal_batch (16 lanes per block, tiles of 4
// blocks).  Within a class the blocks of all its jobs are packed
// into one index space, so a tile may cover many small mip levels and no
// lanes idle at job boundaries.  Each class kernel is persistent: a grid of
// wg_per_cu * CUs workgroups strides over the tiles.  The class launches are
// captured once into a hipGraph, spread over `streams` capture streams so
// that small classes overlap with large ones; a launch of the batch is one
// hipGraphLaunch.
#include "s2tc_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

namespace s2tc {

struct ClassPlan {
	bool normal = false;         // 16-lanes-per-block kernel
	bool all16 = false;          // one-lane-per-block MODE_NORMAL kernel
	int cd = 0, dxt = -1;
	std::vector<JobDesc> jobs;
	std::vector<int> job_index;  // index into the caller's job array
	long long nblocks = 0, texels = 0, bytes = 0;
	int ntiles = 0, tile = 256, pbw_max = 1;
	size_t smem = 0;
	const void *kernel = nullptr;
	int occupancy = 1;           // resident workgroups per CU
	JobDesc *d_jobs = nullptr;
	TileJobs *d_tiles = nullptr;
	double cost = 0;             // relative cost estimate, for ordering
};

}  // namespace s2tc

struct s2tc_amd_batch {
	std::vector<s2tc::ClassPlan> classes;
	int ncu = 256;
	int wg_per_cu = 4, streams = 4;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	std::vector<hipStream_t> cap_streams;
	std::vector<hipEvent_t> cap_events;
};

namespace s2tc {

static int launch_class(const ClassPlan &c, int wg_per_cu, int ncu, hipStream_t st)
{
	long long nb = c.nblocks;
	int ntiles = c.ntiles;
	int grid = wg_per_cu <= 0 ? ntiles : std::min(ntiles, wg_per_cu * ncu);
	if (grid < 1) return 0;
	const JobDesc *jobs = c.d_jobs;
	const TileJobs *tiles = c.d_tiles;
	if (c.normal) {
		int pbw = c.pbw_max;
		void *args[] = {(void *)&jobs, (void *)&tiles, (void *)&nb, (void *)&ntiles, (void *)&pbw};
		return hip_check(hipLaunchKernel(c.kernel, dim3(grid), dim3(64), args, c.smem, st), "launch k_normal_batch");
	}
	void *args[] = {(void *)&jobs, (void *)&tiles, (void *)&nb, (void *)&ntiles};
	return hip_check(hipLaunchKernel(c.kernel, dim3(grid), dim3(256), args, 0, st),
	                 c.all16 ? "launch k_all16_batch" : "launch k_fast_batch");
}

static void destroy_graph(s2tc_amd_batch *b)
{
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	b->exec = nullptr;
	b->graph = nullptr;
}

// Capture all class launches into one graph: classes in decreasing cost are
// dealt round-robin to the capture streams, which fork from and join back
// into the first one.
static int capture(s2tc_amd_batch *b)
{
	destroy_graph(b);
	const int ns = std::max(1, std::min(b->streams, (int)b->cap_streams.size()));
	hipStream_t s0 = b->cap_streams[0];
	if (int r = hip_check(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
		return r;
	int rc = hip_check(hipEventRecord(b->cap_events[0], s0), "hipEventRecord");
	for (int i = 1; i < ns && !rc; ++i)
		rc = hip_check(hipStreamWaitEvent(b->cap_streams[i], b->cap_events[0], 0), "hipStreamWaitEvent");
	for (size_t k = 0; k < b->classes.size() && !rc; ++k)
		rc = launch_class(b->classes[k], b->wg_per_cu, b->ncu, b->cap_streams[k % ns]);
	for (int i = 1; i < ns && !rc; ++i) {
		rc = hip_check(hipEventRecord(b->cap_events[i], b->cap_streams[i]), "hipEventRecord");
		if (!rc) rc = hip_check(hipStreamWaitEvent(s0, b->cap_events[i], 0), "hipStreamWaitEvent");
	}
	hipGraph_t g = nullptr;
	const int rc2 = hip_check(hipStreamEndCapture(s0, &g), "hipStreamEndCapture");
	if (rc) {
		if (g) (void)hipGraphDestroy(g);
		return rc;
	}
	if (rc2) return rc2;
	b->graph = g;
	return hip_check(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
}

static void free_plan(s2tc_amd_batch *b)
{
	destroy_graph(b);
	for (auto &c : b->classes) {
		if (c.d_jobs) (void)hipFree(c.d_jobs);
		c.d_jobs = nullptr;
		c.d_tiles = nullptr;
	}
	for (auto s : b->cap_streams) (void)hipStreamDestroy(s);
	for (auto e : b->cap_events) (void)hipEventDestroy(e);
	b->cap_streams.clear();
	b->cap_events.clear();
}

}  // namespace s2tc

using namespace s2tc;

extern "C" {

int s2tc_amd_batch_create(const s2tc_amd_job *jobs, int njobs, s2tc_amd_batch **out)
{
	if (!out || (njobs > 0 && !jobs) || njobs < 0) return set_error(S2TC_E_ARG, "bad batch arguments");
	*out = nullptr;
	if (int r = ensure_device()) return r;
	s2tc_amd_batch *b = new s2tc_amd_batch;
	int dev = 0;
	(void)hipGetDevice(&dev);
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
		b->ncu = prop.multiProcessorCount;
	// classes: 0..23 quick (cd*3 + dxt), 24..31 MODE_NORMAL with random
	// colours (cd), 32..55 MODE_NORMAL without (32 + cd*3 + dxt)
	std::vector<int> slot(56, -1);
	for (int i = 0; i < njobs; ++i) {
		const s2tc_amd_job &j = jobs[i];
		int dxt = j.dxt, cd = j.cd, refine = j.refine;
		normalize_modes(dxt, cd, refine);
		if (j.width <= 0 || j.height <= 0 || !j.src || !j.dst) {
			delete b;
			return set_error(S2TC_E_ARG, "job %d: bad image", i);
		}
		if (j.nrandom > S2TC_MAX_NRANDOM) {
			delete b;
			return set_error(S2TC_E_RANGE, "job %d: nrandom %d exceeds the GPU limit %d", i, j.nrandom,
			                 S2TC_MAX_NRANDOM);
		}
		if (j.src_kind != S2TC_SRC_QUANTIZED && j.src_kind != S2TC_SRC_RGBA8) {
			delete b;
			return set_error(S2TC_E_ARG, "job %d: bad source kind", i);
		}
		const bool fast = is_fast(cd, j.nrandom);
		const bool all16 = !fast && j.nrandom <= 0;
		const int key = fast ? cd * 3 + dxt : all16 ? 32 + cd * 3 + dxt : 24 + cd;
		if (slot[key] < 0) {
			slot[key] = (int)b->classes.size();
			b->classes.emplace_back();
			ClassPlan &c = b->classes.back();
			c.normal = !fast && !all16;
			c.all16 = all16;
			c.cd = cd;
			c.dxt = c.normal ? -1 : dxt;
			c.tile = c.normal ? 4 : 256;
		}
		ClassPlan &c = b->classes[slot[key]];
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
	int rc = 0;
	for (auto &c : b->classes) {
		const KernelTable &kt = kernel_table(c.cd);
		c.kernel = c.normal ? kt.normal_batch : c.all16 ? kt.all16_batch[c.dxt] : kt.fast_batch[c.dxt];
		if (!c.kernel) {
			rc = set_error(S2TC_E_ARG, "no batch kernel for distance %d", c.cd);
			break;
		}
		c.ntiles = (int)((c.nblocks + c.tile - 1) / c.tile);
		c.smem = c.normal ? (size_t)kt.normal_smem_per_cand * 16 * c.pbw_max : 0;
		// tile -> (first job, last job), by a walk over the packed blocks
		std::vector<TileJobs> tiles((size_t)c.ntiles);
		int jlo = 0;
		const int nj = (int)c.jobs.size();
		for (int t = 0; t < c.ntiles; ++t) {
			const long long b0 = (long long)t * c.tile, b1 = std::min(c.nblocks, b0 + c.tile) - 1;
			while (jlo + 1 < nj && c.jobs[jlo + 1].block0 <= b0) ++jlo;
			int jhi = jlo;
			while (jhi + 1 < nj && c.jobs[jhi + 1].block0 <= b1) ++jhi;
			tiles[(size_t)t] = TileJobs{jlo, jhi};
		}
		const size_t jb = c.jobs.size() * sizeof(JobDesc), tb = tiles.size() * sizeof(TileJobs);
		void *mem = nullptr;
		rc = hip_check(hipMalloc(&mem, jb + tb + 16), "hipMalloc (batch plan)");
		if (rc) break;
		c.d_jobs = (JobDesc *)mem;
		c.d_tiles = (TileJobs *)((uint8_t *)mem + ((jb + 15) & ~(size_t)15));
		rc = hip_check(hipMemcpy(c.d_jobs, c.jobs.data(), jb, hipMemcpyHostToDevice), "hipMemcpy (jobs)");
		if (!rc) rc = hip_check(hipMemcpy(c.d_tiles, tiles.data(), tb, hipMemcpyHostToDevice), "hipMemcpy (tiles)");
		if (rc) break;
		int occ = 1;
		if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, c.kernel, c.normal ? 64 : 256, c.smem) == hipSuccess &&
		    occ > 0)
			c.occupancy = occ;
	}
	if (!rc) {
		std::stable_sort(b->classes.begin(), b->classes.end(),
		                 [](const ClassPlan &x, const ClassPlan &y) { return x.cost > y.cost; });
		for (int i = 0; i < 8 && !rc; ++i) {
			hipStream_t s;
			hipEvent_t e;
			rc = hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
			if (rc) break;
			b->cap_streams.push_back(s);
			rc = hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
			if (rc) break;
			b->cap_events.push_back(e);
		}
	}
	if (!rc) rc = capture(b);
	if (rc) {
		free_plan(b);
		delete b;
		return rc;
	}
	*out = b;
	return 0;
}

int s2tc_amd_batch_set_schedule(s2tc_amd_batch *b, int wg_per_cu, int streams)
{
	if (!b || streams < 1 || streams > 8 || wg_per_cu < 0) return set_error(S2TC_E_ARG, "bad schedule");
	b->wg_per_cu = wg_per_cu;
	b->streams = streams;
	return capture(b);
}

int s2tc_amd_batch_get_schedule(const s2tc_amd_batch *b, int *wg_per_cu, int *streams)
{
	if (!b) return set_error(S2TC_E_ARG, "null batch");
	if (wg_per_cu) *wg_per_cu = b->wg_per_cu;
	if (streams) *streams = b->streams;
	return 0;
}

int s2tc_amd_batch_launch(s2tc_amd_batch *b, void *stream)
{
	if (!b || !b->exec) return set_error(S2TC_E_ARG, "batch not ready");
	return hip_check(hipGraphLaunch(b->exec, (hipStream_t)stream), "hipGraphLaunch");
}

int s2tc_amd_batch_autotune(s2tc_amd_batch *b, void *stream, int reps, double *best_ms)
{
	if (!b) return set_error(S2TC_E_ARG, "null batch");
	if (reps < 1) reps = 1;
	hipStream_t st = (hipStream_t)stream;
	hipEvent_t e0, e1;
	if (int r = hip_check(hipEventCreate(&e0), "hipEventCreate")) return r;
	if (int r = hip_check(hipEventCreate(&e1), "hipEventCreate")) {
		(void)hipEventDestroy(e0);
		return r;
	}
	static const int kWg[] = {1, 2, 4, 8, 16, 0};
	static const int kStreams[] = {1, 2, 4, 8};
	double best = 1e30;
	int bw = b->wg_per_cu, bs = b->streams, rc = 0;
	for (int wg : kWg) {
		for (int ns : kStreams) {
			b->wg_per_cu = wg;
			b->streams = ns;
			if ((rc = capture(b))) break;
			if ((rc = s2tc_amd_batch_launch(b, stream))) break;  // warm-up
			if ((rc = hip_check(hipEventRecord(e0, st), "hipEventRecord"))) break;
			for (int i = 0; i < reps && !rc; ++i) rc = s2tc_amd_batch_launch(b, stream);
			if (!rc) rc = hip_check(hipEventRecord(e1, st), "hipEventRecord");
			if (!rc) rc = hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
			if (rc) break;
			float ms = 0;
			(void)hipEventElapsedTime(&ms, e0, e1);
			const double per = ms / reps;
			if (per < best) {
				best = per;
				bw = wg;
				bs = ns;
			}
		}
		if (rc) break;
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	b->wg_per_cu = bw;
	b->streams = bs;
	if (int r = capture(b)) return r;
	if (rc) return rc;
	if (best_ms) *best_ms = best;
	return 0;
}

int s2tc_amd_batch_class_times(s2tc_amd_batch *b, void *stream, s2tc_amd_class_time *out, int max_out)
{
	if (!b) return set_error(S2TC_E_ARG, "null batch");
	hipStream_t st = (hipStream_t)stream;
	hipEvent_t e0, e1;
	if (int r = hip_check(hipEventCreate(&e0), "hipEventCreate")) return r;
	if (int r = hip_check(hipEventCreate(&e1), "hipEventCreate")) {
		(void)hipEventDestroy(e0);
		return r;
	}
	int rc = 0, n = 0;
	for (const ClassPlan &c : b->classes) {
		rc = launch_class(c, b->wg_per_cu, b->ncu, st);  // warm-up
		if (!rc) rc = hip_check(hipEventRecord(e0, st), "hipEventRecord");
		if (!rc) rc = launch_class(c, b->wg_per_cu, b->ncu, st);
		if (!rc) rc = hip_check(hipEventRecord(e1, st), "hipEventRecord");
		if (!rc) rc = hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
		if (rc) break;
		float ms = 0;
		(void)hipEventElapsedTime(&ms, e0, e1);
		if (out && n < max_out) {
			s2tc_amd_class_time &o = out[n];
			o.normal = c.normal ? 1 : c.all16 ? 2 : 0;
			o.cd = c.cd;
			o.dxt = c.dxt;
			o.njobs = (int)c.jobs.size();
			o.blocks = c.nblocks;
			o.texels = c.texels;
			o.bytes = c.bytes;
			o.ms = ms;
		}
		++n;
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	return rc ? rc : n;
}

int s2tc_amd_batch_destroy(s2tc_amd_batch *b)
{
	if (!b) return 0;
	free_plan(b);
	delete b;
	return 0;
}

int s2tc_amd_encode_jobs_plain(const s2tc_amd_job *jobs, int njobs, void *stream)
{
	if (njobs < 0 || (njobs > 0 && !jobs)) return set_error(S2TC_E_ARG, "bad job list");
	for (int i = 0; i < njobs; ++i) {
		const s2tc_amd_job &j = jobs[i];
		EncodeArgs a{j.src, j.dst, j.src_kind, j.width, j.height, j.dxt, j.cd, j.nrandom, j.refine, j.seed, 0};
		if (int r = launch_encode(a, (hipStream_t)stream)) return r;
	}
	return 0;
}

}  // extern "C"
