/*
 * s2tc_oracle.h - CPU restatement of the reference S2TC encoder.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU library is
 * compared against; it is never linked into, called by or shipped with the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Every function restates a reference function (file:line cited at each
 * definition in s2tc_oracle.c) in plain C, with int32 wrap-around made
 * explicit and float operations in the reference's order.
 *
 * Deviations from the stock reference, both by design:
 *  - rand() is replaced by the counter-based generator of s2tc_prng.h;
 *    bkey is the block key s2tc_prng_block_key(seed, block index).
 *  - the reference reads an uninitialised ca[1] for DXT5 blocks with a single
 *    texel and nrandom <= 0 (s2tc_algorithm.cpp:996-1001, 1006); the oracle
 *    pins it to 0, which is what the reference does when built with
 *    -ftrivial-auto-var-init=zero (oracle/Makefile builds it that way).
 */
#ifndef S2TC_ORACLE_H
#define S2TC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* enum values as in s2tc_algorithm.h:31-63 */
enum { O_DITHER_NONE = 0, O_DITHER_SIMPLE = 1, O_DITHER_FLOYDSTEINBERG = 2 };
enum { O_DXT1 = 0, O_DXT3 = 1, O_DXT5 = 2 };
enum { O_REFINE_NEVER = 0, O_REFINE_ALWAYS = 1, O_REFINE_LOOP = 2 };
enum { O_RGB = 0, O_YUV, O_SRGB, O_SRGB_MIXED, O_AVG, O_WAVG, O_W0AVG, O_NORMALMAP };

/* colour distance between two 5:6:5 colours (r,g,b each 0..31/63/31) */
int oracle_color_dist(int cd, int ar, int ag, int ab, int br, int bg, int bb);
/* srgb_get_y of one colour */
int oracle_srgb_get_y(int r, int g, int b);

/* One 4x4 block from a quantized RGBA buffer (row stride iw texels). */
void oracle_encode_block(uint8_t *out, const uint8_t *rgba, int iw, int w, int h,
                         int dxt, int cd, int nrandom, int refine, uint32_t bkey);

/* rgb565_image: 8-bit RGB(A) -> quantized {r5, g6, b5, a(alphabits)} */
void oracle_rgb565_image(uint8_t *out, const uint8_t *rgba, int w, int h,
                         int srccomps, int alphabits, int dither);

/* tx_compress_dxtn with explicit settings instead of environment variables.
 * fmt is a GL enum (0x83F0..0x83F3).  Returns 0, or -1 for a bad format. */
int oracle_compress_image(int srccomps, int width, int height, const uint8_t *src,
                          unsigned fmt, uint8_t *dest, int dstRowStride,
                          int cd, int nrandom, int refine, int dither, uint64_t seed);

/* Same block loop, already-quantized input; blocks [row0, row1) of block rows. */
void oracle_encode_image_quantized(const uint8_t *q, int width, int height, int dxt,
                                   int cd, int nrandom, int refine, uint64_t seed,
                                   uint8_t *dest, int row0, int row1);

/* fetch_2d_texel_*: fmt as GL enum; texel gets 4 bytes RGBA. */
void oracle_fetch_texel(unsigned fmt, int srcRowStride, const uint8_t *pix,
                        int i, int j, uint8_t *texel);

/* Total colour distance of an encoding: the sum over texels of
 * ColorDist(texel, colour the block decodes to), the per-texel quantity
 * s2tc_try_encode_block minimises (s2tc_algorithm.cpp:606-616), with the
 * decoded 5:6:5 colour chosen by the fetch_2d_texel rules
 * (s2tc_libtxc_dxtn.cpp:35-140).  DXT1 texels with alpha 0 are skipped.
 * q: quantized {r5, g6, b5, a} texels; blocks in raster order. */
int64_t oracle_colordist_error(const uint8_t *q, int width, int height, int dxt, int cd, const uint8_t *blocks);

/* Image_MipReduce32 to (destwidth, destheight) = (1, 1); in may equal out. */
void oracle_mip_reduce(const uint8_t *in, uint8_t *out, int *width, int *height);

/* s2tc_from_s3tc's per-block remap of an S3TC block to S2TC, in place:
 * dxt O_DXT1 -> convert_dxt1a on the 8-byte block, O_DXT3 -> convert_dxt1 on
 * bytes 8-15, O_DXT5 -> convert_dxt1 on bytes 8-15 and convert_dxt5 on bytes
 * 0-7 (s2tc_from_s3tc.cpp:77-190, 254-263). */
void oracle_from_s3tc_block(uint8_t *block, int dxt);

#ifdef __cplusplus
}
#endif

#endif
