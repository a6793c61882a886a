// s2tc_from_s3tc - command-line S3TC -> S2TC transcoder on the GPU.
//
// Same options, messages and output bytes as the reference tool
// (s2tc_from_s3tc.cpp:192-271): -i input DDS (default stdin), -o output
// (default stdout); the 128-byte DDS header is copied unchanged, every
// whole block after it is remapped (s2tc_amd_from_s3tc), a trailing partial
// block is dropped.  Only DXT1, DXT3 and DXT5 files are accepted.  A header
// shorter than 128 bytes is zero-filled (the reference leaves it unset).
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <vector>

#include "s2tc_amd.h"

namespace {

int usage(const char *me)
{
	fprintf(stderr, "usage:\n%s \n    [-i infile.dds]\n    [-o outfile.dds]\n", me);
	return 1;
}

int gpu_fail(const char *what)
{
	fprintf(stderr, "s2tc_from_s3tc: %s: %s\n", what, s2tc_amd_last_error());
	return 3;
}

}  // namespace

int main(int argc, char **argv)
{
	const char *infile = nullptr, *outfile = nullptr;
	int opt;
	while ((opt = getopt(argc, argv, "i:o:")) != -1) {
		switch (opt) {
		case 'i': infile = optarg; break;
		case 'o': outfile = optarg; break;
		default: return usage(argv[0]);
		}
	}
	FILE *in = infile ? fopen(infile, "rb") : stdin;
	if (!in) {
		printf("opening input failed\n");
		return 2;
	}
	FILE *out = outfile ? fopen(outfile, "wb") : stdout;
	if (!out) {
		printf("opening output failed\n");
		return 2;
	}
	uint8_t hdr[128] = {0};
	if (fread(hdr, 1, sizeof hdr, in) == 0 && ferror(in)) return 2;
	int dxt, bs;
	if (!memcmp(hdr + 84, "DXT1", 4)) dxt = 0, bs = 8;
	else if (!memcmp(hdr + 84, "DXT3", 4)) dxt = 1, bs = 16;
	else if (!memcmp(hdr + 84, "DXT5", 4)) dxt = 2, bs = 16;
	else {
		fprintf(stderr, "Only DXT1, DXT3, DXT5 are supported!\n");
		return 1;
	}
	fwrite(hdr, sizeof hdr, 1, out);
	std::vector<uint8_t> data;
	{
		uint8_t buf[1 << 16];
		size_t n;
		while ((n = fread(buf, 1, sizeof buf, in)) > 0) data.insert(data.end(), buf, buf + n);
	}
	const size_t nblocks = data.size() / bs, bytes = nblocks * bs;
	if (nblocks) {
		void *d = nullptr;
		if (s2tc_amd_malloc(&d, bytes)) return gpu_fail("device allocation");
		int rc = s2tc_amd_memcpy_h2d(d, data.data(), bytes);
		if (!rc) rc = s2tc_amd_from_s3tc(d, d, nblocks, dxt, nullptr);
		if (!rc) rc = s2tc_amd_memcpy_d2h(data.data(), d, bytes);
		s2tc_amd_free(d);
		if (rc) return gpu_fail("transcode");
		fwrite(data.data(), bytes, 1, out);
	}
	if (infile) fclose(in);
	if (outfile) fclose(out);
	return 0;
}
