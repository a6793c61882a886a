/*
 * s2tc_algorithm.h - the reference's algorithm-level C API
 * (s2tc_algorithm.h:31-66 of the reference), same enum values and entry
 * points, implemented on the GPU.
 *
 * rgb565_image (reference :38, s2tc_algorithm.cpp:1453-1465): 8-bit RGB(A)
 *   host buffers in and out; the quantization and dithering run on the GPU.
 *
 * s2tc_encode_block_func (reference :65-66, s2tc_algorithm.cpp:1164-1194):
 *   returns a per-4x4-block encoder for (dxt, cd, nrandom, refine) with the
 *   reference's selection rules (quick mode iff nrandom < 0 and cd is not
 *   NORMALMAP; out-of-range values fall back as the reference's do).  Each
 *   call encodes ONE block synchronously on the GPU, which is correct but
 *   slow; whole images go through tx_compress_dxtn or s2tc_amd.h.
 *   For nrandom > 0 each call draws from the counter-based generator keyed
 *   by (S2TC_RANDOM_SEED, per-thread block counter); the counter starts at 0
 *   in every thread and advances by one per call, and
 *   s2tc_amd_set_block_index() (s2tc_amd.h) sets it.
 */
#ifndef S2TC_COMPRESSOR_H
#define S2TC_COMPRESSOR_H

#ifndef S2TC_API
#if defined(__GNUC__)
#define S2TC_API __attribute__((visibility("default")))
#else
#define S2TC_API
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum DitherMode { DITHER_NONE, DITHER_SIMPLE, DITHER_FLOYDSTEINBERG };

S2TC_API void rgb565_image(unsigned char *out, const unsigned char *rgba, int w, int h, int srccomps, int alphabits,
                  enum DitherMode dither);

enum DxtMode { DXT1, DXT3, DXT5 };
enum RefinementMode { REFINE_NEVER, REFINE_ALWAYS, REFINE_LOOP };

typedef enum { RGB, YUV, SRGB, SRGB_MIXED, AVG, WAVG, W0AVG, NORMALMAP } ColorDistMode;

typedef void (*s2tc_encode_block_func_t)(unsigned char *out, const unsigned char *rgba, int iw, int w, int h,
                                         int nrandom);
S2TC_API s2tc_encode_block_func_t s2tc_encode_block_func(enum DxtMode dxt, ColorDistMode cd, int nrandom,
                                                enum RefinementMode refine);

#ifdef __cplusplus
}
#endif

#endif
