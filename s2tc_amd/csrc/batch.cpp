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
#include "plan.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

namespace s2tc {

}  // namespace s2tc

struct s2tc_amd_batch {
	std::vector<s2tc::ClassPlan> classes;
	int ncu = 256;
	int wg_per_cu = 4, streams = 4;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	std::vector<hipStream_t> cap_streams;
	std::vector<hipEvent_t> cap_events;
	void *d_desc = nullptr;  // device copy of every class's job and tile descriptors
	void *d_scratch = nullptr;  // group memory of wide random-candidate classes
};

namespace s2tc {

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
		rc = plan_launch_class(b->classes[k], b->wg_per_cu, b->ncu, b->cap_streams[k % ns]);
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
	if (b->d_desc) (void)hipFree(b->d_desc);
	b->d_desc = nullptr;
	if (b->d_scratch) (void)hipFree(b->d_scratch);
	b->d_scratch = nullptr;
	for (auto s : b->cap_streams) (void)hipStreamDestroy(s);
	for (auto e : b->cap_events) (void)hipEventDestroy(e);
	b->cap_streams.clear();
	b->cap_events.clear();
}

}  // namespace s2tc

namespace s2tc {

// Whole-graph timing of the batch's current plan: capture, one warm-up
// launch, then reps launches between two events.
struct GraphTimer {
	s2tc_amd_batch *b;
	hipStream_t st;
	int reps;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	int init()
	{
		if (int r = hip_check(hipEventCreate(&e0), "hipEventCreate")) return r;
		return hip_check(hipEventCreate(&e1), "hipEventCreate");
	}
	~GraphTimer()
	{
		if (e0) (void)hipEventDestroy(e0);
		if (e1) (void)hipEventDestroy(e1);
	}
	int time(double &per)
	{
		if (int r = capture(b)) return r;
		if (int r = s2tc_amd_batch_launch(b, st)) return r;  // warm-up
		if (int r = hip_check(hipEventRecord(e0, st), "hipEventRecord")) return r;
		for (int i = 0; i < reps; ++i)
			if (int r = s2tc_amd_batch_launch(b, st)) return r;
		if (int r = hip_check(hipEventRecord(e1, st), "hipEventRecord")) return r;
		if (int r = hip_check(hipEventSynchronize(e1), "hipEventSynchronize")) return r;
		float ms = 0;
		(void)hipEventElapsedTime(&ms, e0, e1);
		per = ms / reps;
		return 0;
	}
};

// The schedule grid: workgroups per CU x capture streams; keeps the fastest.
static int tune_schedule(s2tc_amd_batch *b, GraphTimer &gt, double &best)
{
	static const int kWg[] = {1, 2, 4, 8, 16, 0};
	static const int kStreams[] = {1, 2, 4, 8};
	int bw = b->wg_per_cu, bs = b->streams;
	for (int wg : kWg)
		for (int ns : kStreams) {
			b->wg_per_cu = wg;
			b->streams = ns;
			double per = 0;
			if (int r = gt.time(per)) return r;
			if (per < best) {
				best = per;
				bw = wg;
				bs = ns;
			}
		}
	b->wg_per_cu = bw;
	b->streams = bs;
	return 0;
}

// The kernel forms of each all-16 class, one class at a time in decreasing
// cost with the others fixed; keeps a form only if the whole graph gets
// faster.
static int tune_forms(s2tc_amd_batch *b, GraphTimer &gt, double &best)
{
	for (ClassPlan &c : b->classes) {
		if (!c.all16) continue;
		int keep = c.form;
		for (int f = 0; f < 3; ++f) {
			if (f == keep || !c.forms[f]) continue;
			c.form = f;
			double per = 0;
			if (int r = gt.time(per)) {
				c.form = keep;
				return r;
			}
			if (per < best) {
				best = per;
				keep = f;
			}
		}
		c.form = keep;
	}
	return 0;
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
	b->ncu = device_cu_count();
	int rc = plan_classes(jobs, njobs, b->classes, true);
	if (!rc) {
		const size_t bytes = plan_descriptor_bytes(b->classes);
		std::vector<uint8_t> host(bytes);
		rc = hip_check(hipMalloc(&b->d_desc, bytes ? bytes : 16), "hipMalloc (batch plan)");
		if (!rc) {
			plan_pack(b->classes, host.data(), (uint8_t *)b->d_desc);
			rc = hip_check(hipMemcpy(b->d_desc, host.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy (plan)");
		}
		const size_t sbytes = plan_scratch_bytes(b->classes);
		if (!rc && sbytes) {
			rc = hip_check(hipMalloc(&b->d_scratch, sbytes), "hipMalloc (batch scratch)");
			if (!rc) plan_set_scratch(b->classes, (uint8_t *)b->d_scratch);
		}
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
	GraphTimer gt{b, (hipStream_t)stream, reps < 1 ? 1 : reps};
	int rc = gt.init();
	double best = 1e30;
	if (!rc) rc = gt.time(best);  // the current plan is the first candidate
	if (!rc) rc = tune_schedule(b, gt, best);
	if (!rc) rc = tune_forms(b, gt, best);
	if (!rc) rc = tune_schedule(b, gt, best);
	if (int r = capture(b)) return r;
	if (rc) return rc;
	if (best_ms) *best_ms = best;
	return 0;
}

int s2tc_amd_batch_set_form(s2tc_amd_batch *b, int klass, int form)
{
	if (!b || klass < 0 || klass >= (int)b->classes.size()) return set_error(S2TC_E_ARG, "bad class index");
	ClassPlan &c = b->classes[(size_t)klass];
	const int native = c.normal ? FORM_ROW16 : FORM_LANE;
	if (form < 0 || form > 2 || (c.all16 ? !c.forms[form] : form != native))
		return set_error(S2TC_E_ARG, "class %d has no kernel form %d", klass, form);
	if (c.all16) c.form = form;
	return capture(b);
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
		rc = plan_launch_class(c, b->wg_per_cu, b->ncu, st);  // warm-up
		if (!rc) rc = hip_check(hipEventRecord(e0, st), "hipEventRecord");
		if (!rc) rc = plan_launch_class(c, b->wg_per_cu, b->ncu, st);
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
			o.refine = c.refine;
			o.njobs = (int)c.jobs.size();
			o.blocks = c.nblocks;
			o.texels = c.texels;
			o.bytes = c.bytes;
			o.ms = ms;
			o.form = c.all16 ? c.form : c.normal ? FORM_ROW16 : FORM_LANE;
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
