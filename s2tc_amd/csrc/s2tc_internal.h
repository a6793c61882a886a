// s2tc_internal.h - shared declarations of the host-side library code.
#pragma once

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__ 1
#endif
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

#include "kernel_table.h"
#include "s2tc_amd.h"

namespace s2tc {

enum { DXT1 = 0, DXT3 = 1, DXT5 = 2 };
enum { REFINE_NEVER = 0, REFINE_ALWAYS = 1, REFINE_LOOP = 2 };
enum { CD_RGB = 0, CD_YUV = 1, CD_SRGB = 2, CD_SRGB_MIXED = 3, CD_AVG = 4, CD_WAVG = 5, CD_W0AVG = 6,
       CD_NORMALMAP = 7 };
enum { DITHER_NONE = 0, DITHER_SIMPLE = 1, DITHER_FLOYDSTEINBERG = 2 };

struct EncodeArgs {
	const void *src;
	void *dst;
	int src_kind;
	int width, height;
	int dxt, cd, nrandom, refine;
	uint64_t seed;
	uint32_t block_index_base;  // generator block index of the image's first block
};

int set_error(int code, const char *fmt, ...);
int hip_check(hipError_t e, const char *what);
int ensure_device();
const KernelTable &kernel_table(int cd);
void normalize_modes(int &dxt, int &cd, int &refine);
bool is_fast(int cd, int nrandom);
int normal_pbw(int nrandom);
// Encodes one image on `stream`.  Random-candidate encodes above
// S2TC_LDS_NRANDOM need device scratch: from `ext_scratch` (ext_bytes long;
// the grid shrinks to what fits) when given, else allocated stream-ordered
// around the launch.
int launch_encode(const EncodeArgs &a, hipStream_t stream, void *ext_scratch = nullptr, size_t ext_bytes = 0);

// rgb565_image on the device (dither.hip); d_scratch holds rgb565_scratch_bytes(w, h).
// d_carry (4 bytes, NONE and SIMPLE only): the image is a band of rows of a
// larger one; SIMPLE starts from the error state in d_carry when carry_in is
// set (else from the reference's initial state) and leaves its end state
// there for the next band.
size_t rgb565_scratch_bytes(int w, int h);
int launch_rgb565(const void *d_src, int srccomps, int w, int h, int alphabits, int dither, void *d_out,
                  void *d_scratch, hipStream_t st, void *d_carry = nullptr, int carry_in = 0);

inline int block_bytes(int dxt) { return dxt == DXT1 ? 8 : 16; }

}  // namespace s2tc

