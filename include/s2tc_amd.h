/*
 * s2tc_amd.h - native (GPU) entry points of the MI355X S2TC encoder.
 *
 * These have no counterpart in the reference; they expose the GPU work at
 * the granularity it is efficient at (whole images, batches, a serving
 * loop).  The reference-compatible entry points are in txc_dxtn.h and
 * s2tc_algorithm.h.  All functions return 0 on success and a negative
 * S2TC_E* code on failure; s2tc_amd_last_error() gives the message.
 *
 * Enum values are those of s2tc_algorithm.h: DxtMode DXT1=0 DXT3=1 DXT5=2;
 * ColorDistMode RGB=0 YUV=1 SRGB=2 SRGB_MIXED=3 AVG=4 WAVG=5 W0AVG=6
 * NORMALMAP=7; RefinementMode NEVER=0 ALWAYS=1 LOOP=2; DitherMode NONE=0
 * SIMPLE=1 FLOYDSTEINBERG=2.
 *
 * Random candidates (nrandom > 0) come from the counter-based generator in
 * s2tc_prng.h keyed by (seed, raster block index within the image).
 */
#ifndef S2TC_AMD_H
#define S2TC_AMD_H

#include <stddef.h>
#include <stdint.h>

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

#define S2TC_E_ARG      (-1)   /* invalid argument */
#define S2TC_E_HIP      (-2)   /* HIP runtime error */
#define S2TC_E_NOGPU    (-3)   /* no usable GPU */
#define S2TC_E_RANGE    (-4)   /* a size limit of the GPU path is exceeded */
#define S2TC_E_BUDGET   (-5)   /* request exceeds the serving memory budget */

/* largest nrandom the GPU path supports: the candidate list (16 + nrandom
 * colours, rounded up to 16) and its distance rows for four blocks must fit
 * the 64 KiB of LDS a workgroup can allocate */
#define S2TC_MAX_NRANDOM 144

/* source kinds for s2tc_amd_encode_image */
#define S2TC_SRC_QUANTIZED 0   /* rgb565_image output: r5 g6 b5 a, 4 bytes per texel */
#define S2TC_SRC_RGBA8     1   /* 8-bit RGBA, quantized on load with DITHER_NONE */

S2TC_API const char *s2tc_amd_last_error(void);
S2TC_API int s2tc_amd_device_count(void);

/* device memory and stream plumbing (HIP device 0 unless set otherwise) */
S2TC_API int s2tc_amd_malloc(void **dptr, size_t bytes);
S2TC_API int s2tc_amd_free(void *dptr);
S2TC_API int s2tc_amd_memcpy_h2d(void *dst, const void *src, size_t bytes);
S2TC_API int s2tc_amd_memcpy_d2h(void *dst, const void *src, size_t bytes);
S2TC_API int s2tc_amd_memset(void *dst, int value, size_t bytes);
S2TC_API int s2tc_amd_synchronize(void);

/* Encode one image already resident on the device.  d_src: width*height
 * texels of 4 bytes (row stride = width); d_dst: ceil(w/4)*ceil(h/4) blocks
 * of 8 (DXT1) or 16 bytes in raster order.  Asynchronous on stream (a
 * hipStream_t, NULL = default stream). */
S2TC_API int s2tc_amd_encode_image(const void *d_src, int src_kind, int width, int height, void *d_dst,
                          int dxt, int cd, int nrandom, int refine, uint64_t seed, void *stream);

#ifdef __cplusplus
}
#endif

#endif
