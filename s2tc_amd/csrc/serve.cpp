// serve.cpp - continuous-batching request loop (s2tc_amd_serve).
//
// Requests (each a texture: a list of images, e.g. its mip levels) arrive
// at given times.  One host thread runs the loop:
//   - arrivals join a FIFO queue;
//   - a request is admitted when the bytes of admitted, unfinished requests
//     (its source bytes plus its output bytes) stay within the HBM budget and
//     its output fits in the device output arena (first-fit allocator);
//   - whenever one of `streams` batch slots is free, all admissible queued
//     requests (up to max_batch) form one batch: its images are grouped into
//     kernel classes (plan.h), the class descriptors go to the device with
//     one copy, and the class kernels run on the slot's stream, so batches of
//     different slots overlap on the GPU;
//   - a batch's requests complete when its event has fired; completion frees
//     their arena space and budget.
// Latency = completion time - arrival time.  With serial_plain the loop
// instead runs one request at a time with one plain launch per image (the
// per-texture baseline).
#include "plan.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

namespace s2tc {

// First-fit allocator over [0, size) with coalescing frees.
class Arena {
public:
	explicit Arena(size_t size) { free_[0] = size; }
	bool alloc(size_t n, size_t &off)
	{
		for (auto it = free_.begin(); it != free_.end(); ++it) {
			if (it->second >= n) {
				off = it->first;
				const size_t rest = it->second - n;
				const size_t roff = it->first + n;
				free_.erase(it);
				if (rest) free_[roff] = rest;
				return true;
			}
		}
		return false;
	}
	void release(size_t off, size_t n)
	{
		auto it = free_.emplace(off, n).first;
		auto nx = std::next(it);
		if (nx != free_.end() && it->first + it->second == nx->first) {
			it->second += nx->second;
			free_.erase(nx);
		}
		if (it != free_.begin()) {
			auto pv = std::prev(it);
			if (pv->first + pv->second == it->first) {
				pv->second += it->second;
				free_.erase(it);
			}
		}
	}

private:
	std::map<size_t, size_t> free_;
};

struct Slot {
	hipStream_t st = nullptr;
	hipEvent_t ev = nullptr;
	void *d_desc = nullptr;
	size_t d_cap = 0;
	void *d_scratch = nullptr;  // group memory of wide random-candidate classes
	size_t s_cap = 0;
	uint8_t *h_desc = nullptr;
	size_t h_cap = 0;
	std::vector<int> reqs;
	bool busy = false;
};

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

static int grow_slot(Slot &s, size_t bytes)
{
	if (bytes <= s.d_cap) return 0;
	bytes = std::max(bytes * 2, (size_t)1 << 20);
	if (s.d_desc) (void)hipFree(s.d_desc);
	if (s.h_desc) (void)hipHostFree(s.h_desc);
	s.d_desc = nullptr;
	s.h_desc = nullptr;
	s.d_cap = s.h_cap = 0;
	if (int r = hip_check(hipMalloc(&s.d_desc, bytes), "hipMalloc (serve descriptors)")) return r;
	if (int r = hip_check(hipHostMalloc((void **)&s.h_desc, bytes, hipHostMallocDefault), "hipHostMalloc")) return r;
	s.d_cap = s.h_cap = bytes;
	return 0;
}

static int grow_scratch(Slot &s, size_t bytes)
{
	if (bytes <= s.s_cap) return 0;
	if (s.d_scratch) (void)hipFree(s.d_scratch);
	s.d_scratch = nullptr;
	s.s_cap = 0;
	if (int r = hip_check(hipMalloc(&s.d_scratch, bytes), "hipMalloc (serve scratch)")) return r;
	s.s_cap = bytes;
	return 0;
}

}  // namespace s2tc

using namespace s2tc;

extern "C" int s2tc_amd_serve(const s2tc_amd_job *jobs, int njobs, const s2tc_amd_request *reqs, int nreq,
                              const s2tc_amd_serve_opts *opts, double *latency_ms, s2tc_amd_serve_stats *stats)
{
	if (!jobs || !reqs || !opts || nreq < 0 || njobs < 0) return set_error(S2TC_E_ARG, "bad serve arguments");
	if (int r = ensure_device()) return r;
	const bool serial = opts->serial_plain != 0;
	const int nslots = serial ? 1 : std::max(1, std::min(16, opts->streams));
	const int max_batch = serial ? 1 : std::max(1, opts->max_batch);
	const size_t budget = opts->hbm_budget;
	const int ncu = device_cu_count();
	for (int i = 0; i < nreq; ++i) {
		if (reqs[i].first_job < 0 || reqs[i].njobs < 1 || reqs[i].first_job + reqs[i].njobs > njobs)
			return set_error(S2TC_E_ARG, "request %d: bad job range", i);
		if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s)
			return set_error(S2TC_E_ARG, "requests must be sorted by arrival time");
	}
	// per-request bytes: outputs (arena) and sources (budget only)
	std::vector<size_t> out_bytes((size_t)nreq, 0), need((size_t)nreq, 0);
	size_t max_out = 0, total_out = 0;
	for (int i = 0; i < nreq; ++i) {
		size_t ob = 0, sb = 0;
		for (int k = 0; k < reqs[i].njobs; ++k) {
			const s2tc_amd_job &j = jobs[reqs[i].first_job + k];
			int dxt = j.dxt, cd = j.cd, rf = j.refine;
			normalize_modes(dxt, cd, rf);
			ob += align256((size_t)((j.width + 3) / 4) * ((j.height + 3) / 4) * block_bytes(dxt));
			sb += (size_t)j.width * j.height * 4;
		}
		out_bytes[(size_t)i] = ob;
		need[(size_t)i] = ob + sb;
		max_out = std::max(max_out, ob);
		total_out += ob;
	}
	const size_t arena_bytes = std::max<size_t>(256, std::min(budget, total_out));
	void *arena_mem = nullptr;
	if (int r = hip_check(hipMalloc(&arena_mem, arena_bytes), "hipMalloc (serve arena)")) return r;
	Arena arena(arena_bytes);
	std::vector<Slot> slots((size_t)nslots);
	int rc = 0;
	for (auto &s : slots) {
		if (!rc) rc = hip_check(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking), "hipStreamCreate");
		if (!rc) rc = hip_check(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming), "hipEventCreate");
	}
	std::vector<size_t> out_off((size_t)nreq, 0);
	std::vector<double> lat((size_t)nreq, -1.0);
	std::deque<int> queue;
	size_t used = 0;
	int next = 0, done = 0, rejected = 0, batches = 0, inflight = 0, max_queue = 0, max_inflight = 0;
	long long texels = 0;
	std::vector<s2tc_amd_job> bjobs;
	std::vector<ClassPlan> classes;
	typedef std::chrono::steady_clock clk;
	const auto t0 = clk::now();
	auto now_s = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
	double t_last = 0;

	while (!rc && done + rejected < nreq) {
		bool progress = false;
		double now = now_s();
		while (next < nreq && reqs[next].arrival_s <= now) {
			if (need[(size_t)next] > budget || out_bytes[(size_t)next] > arena_bytes) {
				lat[(size_t)next] = -1.0;  // can never be admitted
				++rejected;
			} else {
				queue.push_back(next);
			}
			++next;
			progress = true;
		}
		max_queue = std::max(max_queue, (int)queue.size());
		// completions
		for (auto &s : slots) {
			if (!s.busy) continue;
			const hipError_t q = hipEventQuery(s.ev);
			if (q == hipErrorNotReady) continue;
			if (q != hipSuccess) {
				rc = hip_check(q, "hipEventQuery");
				break;
			}
			now = now_s();
			for (int r : s.reqs) {
				lat[(size_t)r] = (now - reqs[r].arrival_s) * 1e3;
				arena.release(out_off[(size_t)r], out_bytes[(size_t)r]);
				used -= need[(size_t)r];
				for (int k = 0; k < reqs[r].njobs; ++k) {
					const s2tc_amd_job &j = jobs[reqs[r].first_job + k];
					texels += (long long)j.width * j.height;
				}
				++done;
				--inflight;
			}
			t_last = now;
			s.reqs.clear();
			s.busy = false;
			progress = true;
		}
		if (rc) break;
		// dispatch
		for (auto &s : slots) {
			if (s.busy || queue.empty()) continue;
			bjobs.clear();
			while (!queue.empty() && (int)s.reqs.size() < max_batch) {
				const int r = queue.front();
				size_t off;
				if (used + need[(size_t)r] > budget || !arena.alloc(out_bytes[(size_t)r], off)) break;
				queue.pop_front();
				used += need[(size_t)r];
				out_off[(size_t)r] = off;
				s.reqs.push_back(r);
				size_t o = off;
				for (int k = 0; k < reqs[r].njobs; ++k) {
					s2tc_amd_job j = jobs[reqs[r].first_job + k];
					int dxt = j.dxt, cd = j.cd, rf = j.refine;
					normalize_modes(dxt, cd, rf);
					j.dst = (uint8_t *)arena_mem + o;
					o += align256((size_t)((j.width + 3) / 4) * ((j.height + 3) / 4) * block_bytes(dxt));
					bjobs.push_back(j);
				}
			}
			if (s.reqs.empty()) break;  // the head of the queue does not fit yet
			inflight += (int)s.reqs.size();
			max_inflight = std::max(max_inflight, inflight);
			if (serial) {
				for (const auto &j : bjobs) {
					EncodeArgs a{j.src, j.dst, j.src_kind, j.width, j.height, j.dxt, j.cd, j.nrandom, j.refine, j.seed, 0};
					if ((rc = launch_encode(a, s.st))) break;
				}
			} else {
				rc = plan_classes(bjobs.data(), (int)bjobs.size(), classes);
				const size_t db = rc ? 0 : plan_descriptor_bytes(classes);
				if (!rc) rc = grow_slot(s, db);
				const size_t sb = rc ? 0 : plan_scratch_bytes(classes);
				if (!rc && sb) rc = grow_scratch(s, sb);
				if (!rc) {
					plan_set_scratch(classes, (uint8_t *)s.d_scratch);
					plan_pack(classes, s.h_desc, (uint8_t *)s.d_desc);
					rc = hip_check(hipMemcpyAsync(s.d_desc, s.h_desc, db, hipMemcpyHostToDevice, s.st), "H2D (descriptors)");
				}
				for (size_t c = 0; c < classes.size() && !rc; ++c)
					rc = plan_launch_class(classes[c], opts->wg_per_cu, ncu, s.st);
			}
			if (!rc && opts->keep_outputs) {
				size_t k = 0;
				for (int r : s.reqs)
					for (int q = 0; q < reqs[r].njobs && !rc; ++q, ++k) {
						const s2tc_amd_job &j = bjobs[k];
						int dxt = j.dxt, cd = j.cd, rf = j.refine;
						normalize_modes(dxt, cd, rf);
						const size_t n = (size_t)((j.width + 3) / 4) * ((j.height + 3) / 4) * block_bytes(dxt);
						void *host = opts->keep_outputs[reqs[r].first_job + q];
						if (host) rc = hip_check(hipMemcpyAsync(host, j.dst, n, hipMemcpyDeviceToHost, s.st), "D2H (outputs)");
					}
			}
			if (!rc) rc = hip_check(hipEventRecord(s.ev, s.st), "hipEventRecord");
			if (rc) break;
			s.busy = true;
			++batches;
			progress = true;
		}
		if (rc) break;
		if (!progress) {
			// nothing to do: wait a little (until the next arrival at most)
			const double wait = next < nreq ? reqs[next].arrival_s - now_s() : 1e-5;
			if (wait > 2e-5) std::this_thread::sleep_for(std::chrono::microseconds((long)std::min(wait * 1e6 - 10, 50.0)));
			else std::this_thread::yield();
		}
	}
	for (auto &s : slots) {
		if (s.busy) (void)hipEventSynchronize(s.ev);
		if (s.ev) (void)hipEventDestroy(s.ev);
		if (s.st) (void)hipStreamDestroy(s.st);
		if (s.d_desc) (void)hipFree(s.d_desc);
		if (s.d_scratch) (void)hipFree(s.d_scratch);
		if (s.h_desc) (void)hipHostFree(s.h_desc);
	}
	(void)hipFree(arena_mem);
	if (rc) return rc;
	if (latency_ms)
		for (int i = 0; i < nreq; ++i) latency_ms[i] = lat[(size_t)i];
	if (stats) {
		std::vector<double> v;
		for (double x : lat)
			if (x >= 0) v.push_back(x);
		std::sort(v.begin(), v.end());
		auto pct = [&](double p) {
			if (v.empty()) return 0.0;
			size_t k = (size_t)std::min((double)v.size() - 1, std::ceil(p * v.size()) - 1);
			return v[k];
		};
		double sum = 0;
		for (double x : v) sum += x;
		memset(stats, 0, sizeof *stats);
		stats->completed = done;
		stats->rejected = rejected;
		stats->batches = batches;
		stats->max_queue = max_queue;
		stats->max_inflight = max_inflight;
		stats->wall_s = t_last;
		stats->texels = texels;
		stats->p50_ms = pct(0.50);
		stats->p99_ms = pct(0.99);
		stats->mean_ms = v.empty() ? 0 : sum / v.size();
		stats->max_ms = v.empty() ? 0 : v.back();
	}
	return 0;
}
