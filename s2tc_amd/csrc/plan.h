// plan.h - grouping of images into kernel classes for the batch and serving paths.
#pragma once

#include <vector>

#include "s2tc_internal.h"

namespace s2tc {

// One kernel class of a batch: the jobs that run the same kernel.
//   quick mode (nrandom < 0, distance != NORMALMAP): one class per
//     (distance, format, refinement), k_fast_batch, one lane per block,
//     tiles of 256;
//   MODE_NORMAL without random colours: one class per (distance, format,
//     refinement), k_all16_batch, one lane per block, tiles of 256;
//   MODE_NORMAL with random colours: one class per distance (formats mixed),
//     k_normal_batch, 16 lanes per block, tiles of 4 blocks; jobs with more
//     than S2TC_LDS_NRANDOM random colours form their own class per distance
//     (k_normal_batch_wide, group memory in device scratch).
// Kernel forms of a class (s2tc_amd_batch_set_form): quick and
// random-candidate classes have one; all-16 classes have up to three:
enum {
	FORM_LANE = 0,        // one lane per block, one wave per SIMD (k_all16_batch)
	FORM_LANE_2WAVES = 1, // one lane per block, two waves per SIMD (k_all16_batch_alone)
	FORM_ROW16 = 2,       // 16 lanes (one DPP row) per block, 4 blocks per wave (k_normal_batch)
};

struct ClassPlan {
	bool normal = false;         // 16-lanes-per-block kernel
	bool wide = false;           // ... with its group memory in device scratch
	bool all16 = false;          // one-lane-per-block MODE_NORMAL kernel
	int cd = 0, dxt = -1;
	int refine = -1;             // one-lane classes: the refinement all their jobs use
	std::vector<JobDesc> jobs;
	std::vector<TileJobs> tiles;
	std::vector<int> job_index;  // index into the caller's job array
	long long nblocks = 0, texels = 0, bytes = 0;
	int ntiles = 0, tile = 256, pbw_max = 1;
	size_t smem = 0;
	size_t scratch = 0;          // wide classes: device scratch bytes (wide_grid slices)
	int wide_grid = 0;           // wide classes: workgroups of every launch
	uint8_t *d_scratch = nullptr;
	const void *kernel = nullptr;
	// all-16 classes: the kernel of each form (null where the form does not
	// exist) and the tiles of 4 blocks the 16-lane form strides over
	const void *forms[3] = {nullptr, nullptr, nullptr};
	int form = FORM_LANE;
	std::vector<TileJobs> tiles4;
	int ntiles4 = 0;
	size_t smem4 = 0;
	JobDesc *d_jobs = nullptr;   // device addresses (set by plan_pack)
	TileJobs *d_tiles = nullptr;
	TileJobs *d_tiles4 = nullptr;
	double cost = 0;             // relative cost estimate, for ordering
};

int device_cu_count();
// Groups jobs into classes (sorted by decreasing cost) with their job and
// tile descriptors on the host; validates every job.  with_forms: also the
// descriptors every kernel form of the all-16 classes needs (a batch that
// may be tuned); otherwise each class has only its default form.
int plan_classes(const s2tc_amd_job *jobs, int njobs, std::vector<ClassPlan> &classes, bool with_forms = false);
// Bytes of device memory that hold all descriptors of the classes.
size_t plan_descriptor_bytes(const std::vector<ClassPlan> &classes);
// Lays the descriptors out for a device buffer at `dev`: writes them to the
// host staging buffer `host` (same layout) and sets d_jobs / d_tiles.
void plan_pack(std::vector<ClassPlan> &classes, uint8_t *host, uint8_t *dev);
// Bytes of device scratch memory the classes need (wide classes only), and
// its assignment to the classes from a buffer at `dev`.
size_t plan_scratch_bytes(const std::vector<ClassPlan> &classes);
void plan_set_scratch(std::vector<ClassPlan> &classes, uint8_t *dev);
// Workgroups and scratch bytes of a wide random-candidate launch over
// `ntiles` tiles with `pbw` row blocks: slices of 4 * group memory, at most
// 4 workgroups per CU and about 1 GiB in total.
void wide_launch_shape(const KernelTable &kt, int pbw, long long ntiles, int ncu, int &grid, size_t &bytes);
// One launch of a class kernel: a grid of wg_per_cu * ncu workgroups (all
// tiles when wg_per_cu is 0) striding over the class's tiles.
int plan_launch_class(const ClassPlan &c, int wg_per_cu, int ncu, hipStream_t st);

}  // namespace s2tc
