/* txc_capture.c - test infrastructure: a libtxc_dxtn stand-in that records
 * what the reference CLI passes to tx_compress_dxtn.
 *
 * Loaded into the reference's own s2tc_compress (runtime-linking build) with
 * -l, it appends every call's (width, height, RGBA8 texels) to the file named
 * by S2TC_CAPTURE_FILE and writes zero blocks.  The CLI calls it once per mip
 * level after Image_MipReduce32 (s2tc_compress.c:722-733), so the file holds
 * the reference's own mip chain, against which the oracle's restatement and
 * the GPU mip kernel are checked.  File format per call: int32 width, int32
 * height, int32 srccomps, then width*height*srccomps bytes.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void tx_compress_dxtn(int srccomps, int width, int height, const unsigned char *src, unsigned int fmt,
                      unsigned char *dst, int dst_row_stride)
{
	const char *path = getenv("S2TC_CAPTURE_FILE");
	int bs = (fmt == 0x83F0 || fmt == 0x83F1) ? 8 : 16;
	int bw = (width + 3) / 4, bh = (height + 3) / 4, y;
	for (y = 0; y < bh; ++y) memset(dst + (size_t)y * dst_row_stride, 0, (size_t)bw * bs);
	if (path) {
		FILE *f = fopen(path, "ab");
		if (f) {
			int32_t hdr[3] = {width, height, srccomps};
			fwrite(hdr, sizeof hdr, 1, f);
			fwrite(src, (size_t)width * height * srccomps, 1, f);
			fclose(f);
		}
	}
}
