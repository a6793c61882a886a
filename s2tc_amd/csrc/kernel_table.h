// kernel_table.h - per-distance-mode table of encoder kernels (host side).
#pragma once

struct KernelTable {
	const void *fast[3][2];       // [dxt][raw]: k_fast<CD, dxt, raw>, null for NORMALMAP
	const void *normal[2];        // [raw]: k_normal<CD, raw>
	int normal_smem_per_cand;     // LDS bytes per candidate slot for a 4-block workgroup
};

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
