/*
 * txc_dxtn.h - libtxc_dxtn-compatible C ABI of the MI355X S2TC encoder.
 *
 * Same names, argument meaning and enum values as the reference's
 * txc_dxtn.h:38-49, so that s2tc_compress -l, GL drivers and other
 * dlopen() users can load libtxc_dxtn.so from this project in its place.
 *
 * tx_compress_dxtn replaces s2tc_libtxc_dxtn.cpp:142-299:
 *   - configuration from the same environment variables, read on every call:
 *     S2TC_DITHER_MODE (NONE|SIMPLE|FLOYDSTEINBERG, default SIMPLE),
 *     S2TC_COLORDIST_MODE (RGB|YUV|SRGB|SRGB_MIXED|AVG|WAVG|W0AVG|NORMALMAP,
 *     default WAVG), S2TC_RANDOM_COLORS (atoi, default -1),
 *     S2TC_REFINE_COLORS (NEVER|ALWAYS|LOOP, default ALWAYS); invalid values
 *     print the reference's message and keep the default;
 *   - plus S2TC_RANDOM_SEED (strtoull, default 0): the key of the
 *     counter-based generator that replaces rand() for nrandom > 0;
 *   - destformat 0x83F0/0x83F1 -> DXT1, 0x83F2 -> DXT3, 0x83F3 -> DXT5,
 *     anything else prints "libdxtn: Bad dstFormat %d in tx_compress_dxtn"
 *     and returns without writing;
 *   - the reference's destination row padding rule (s2tc_libtxc_dxtn.cpp:243,
 *     :261, :279);
 *   - synchronous: dest is complete on return.  Callable from several host
 *     threads at once (each call uses its own stream and device buffers).
 * The quantization and the block encoding run on the GPU.  A GPU error is
 * reported on stderr and leaves dest unwritten.
 *
 * fetch_2d_texel_* replace s2tc_libtxc_dxtn.cpp:35-140 (S2TC decoders: index
 * 2/3 render as a c0/c1 checkerboard).  They decode one texel on the host.
 */
#ifndef _TXC_DXTN_H
#define _TXC_DXTN_H

#if defined(__has_include)
#if __has_include(<GL/gl.h>)
#include <GL/gl.h>
#define S2TC_HAVE_GL_TYPES 1
#endif
#endif
#ifndef S2TC_HAVE_GL_TYPES
typedef int GLint;
typedef unsigned char GLubyte;
typedef unsigned int GLenum;
typedef void GLvoid;
#endif

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

S2TC_API void fetch_2d_texel_rgb_dxt1(GLint srcRowStride, const GLubyte *pixdata, GLint i, GLint j, GLvoid *texel);
S2TC_API void fetch_2d_texel_rgba_dxt1(GLint srcRowStride, const GLubyte *pixdata, GLint i, GLint j, GLvoid *texel);
S2TC_API void fetch_2d_texel_rgba_dxt3(GLint srcRowStride, const GLubyte *pixdata, GLint i, GLint j, GLvoid *texel);
S2TC_API void fetch_2d_texel_rgba_dxt5(GLint srcRowStride, const GLubyte *pixdata, GLint i, GLint j, GLvoid *texel);

S2TC_API void tx_compress_dxtn(GLint srccomps, GLint width, GLint height, const GLubyte *srcPixData, GLenum destformat,
                      GLubyte *dest, GLint dstRowStride);

#ifdef __cplusplus
}
#endif

#endif /* _TXC_DXTN_H */
