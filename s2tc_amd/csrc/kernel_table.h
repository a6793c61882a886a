// kernel_table.h - per-distance-mode table of encoder kernels (host side).
#pragma once

#include <cstddef>
#include <cstdint>

// One image (a texture's mip level) of a batch, as the persistent kernels see it.
struct JobDesc {
	const uint32_t *src;  // width*height texels, 4 bytes each
	uint8_t *dst;         // blocks in raster order
	int width, height;
	int dxt, nrandom, refine, raw;  // raw: RGBA8 quantized on load with DITHER_NONE
	unsigned long long seed;
	long long block0;     // first block of this job in its class's packed block index space
	uint32_t bw_magic;    // min(2^32 / bw, 2^32 - 1) for bw = block columns: see div_blocks
	uint32_t pad;
};

// The multiplier of div_blocks for bw block columns (bw >= 1).
inline uint32_t bw_magic(int bw)
{
	const unsigned long long m = (1ull << 32) / (unsigned long long)bw;
	return m > 0xffffffffull ? 0xffffffffu : (uint32_t)m;
}

// First and last job index a batch tile touches.
struct TileJobs {
	int lo, hi;
};

struct KernelTable {
	const void *fast[3][2];       // [dxt][raw]: k_fast<CD, dxt, raw>, null for NORMALMAP
	const void *fast_int[3][2][3];  // [dxt][raw][refine]: k_fast_interior (full blocks, 16-byte rows), null for NORMALMAP
	const void *normal[2];        // [raw]: k_normal<CD, raw>
	const void *normal_wide[2];   // [raw]: k_normal_wide<CD, raw> (group memory in global scratch)
	int normal_smem_cand;         // LDS bytes per candidate slot of a 4-block workgroup (candidate words)
	int normal_smem_row;          // LDS bytes per stored distance row of a 4-block workgroup
	int normal_smem_fixed;        // plus these bytes (features of the 16 texel candidates of 4 blocks)
	const void *fast_batch[3][3]; // [dxt][refine]: persistent quick-mode kernel over a job list
	const void *normal_batch;     // persistent MODE_NORMAL kernel over a job list
	const void *normal_batch_wide;  // the same with group memory in global scratch
	const void *all16[3][2];      // [dxt][raw]: MODE_NORMAL without random colours, one lane per block
	const void *all16_batch[3][3];  // [dxt][refine]: its persistent batch form
	const void *all16_batch_alone[3][3];  // the form for a batch of this one class
	const void *error[3][2];      // [dxt][raw]: k_colordist_error (total colour distance of an encoding)
	const void *init;             // per-device table set-up kernel (256 x 256 threads, one int argument), or null
	const void *set_scan_cap;     // k_set_scan_cap (one workgroup of 64 threads, one uint32_t argument)
};

// Row blocks (16 rows each) of a block's distance matrix that the
// random-candidate kernels keep in registers instead of LDS.
#define S2TC_REG_RB 2

// Group memory of a four-block workgroup of the random-candidate kernels
// with pbw row blocks: 16 pbw candidate slots, the rows from 16 S2TC_REG_RB
// on and the texel features.
inline size_t normal_group_smem(const KernelTable &kt, int pbw)
{
	const int stored = pbw > S2TC_REG_RB ? pbw - S2TC_REG_RB : 0;
	return (size_t)kt.normal_smem_cand * 16 * (size_t)pbw + (size_t)kt.normal_smem_row * 16 * (size_t)stored +
	       (size_t)kt.normal_smem_fixed;
}

extern "C" {
void s2tc_kernels_cd0(KernelTable *);
void s2tc_kernels_cd1(KernelTable *);
void s2tc_kernels_cd2(KernelTable *);
void s2tc_kernels_cd3(KernelTable *);
void s2tc_kernels_cd4(KernelTable *);
void s2tc_kernels_cd5(KernelTable *);
void s2tc_kernels_cd6(KernelTable *);
void s2tc_kernels_cd7(KernelTable *);
}
