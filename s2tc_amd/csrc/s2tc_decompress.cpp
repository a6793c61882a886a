// s2tc_decompress - DDS (top mip level) to 32-bit TGA.
//
// Same options, messages and output bytes as the reference tool
// (s2tc_decompress.c:99-214): -i input DDS (default stdin), -o output TGA
// (default stdout), -l the libtxc_dxtn.so to decode with (default: this
// project's, next to the executable).  The output is an 18-byte TGA header
// (type 2, 32 bits, top-left origin) and the top level's texels as B, G, R, A
// in raster order.  With this project's library the whole level is decoded
// on the GPU (s2tc_amd_decode_image, the fetch_2d_texel_* rules); with any
// other library its fetch_2d_texel_* functions decode texel by texel, as
// the reference does.  Blocks missing from a short file read as zero bytes
// (the reference leaves them unset).
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <string>
#include <vector>

namespace {

typedef void fetch_t(int, const unsigned char *, int, int, void *);
typedef int malloc_t(void **, size_t);
typedef int free_t(void *);
typedef int copy_t(void *, const void *, size_t);
typedef int decode_t(const void *, unsigned, int, int, void *, void *);
typedef const char *err_t(void);

int usage(const char *me)
{
	fprintf(stderr, "usage:\n%s \n    [-i infile.tga]\n    [-o outfile.dds]\n    [-l path_to_libtxc_dxtn.so]\n", me);
	return 1;
}

std::string default_library()
{
	char buf[4096];
	ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
	if (n <= 0) return "libtxc_dxtn.so";
	buf[n] = 0;
	std::string s(buf);
	return s.substr(0, s.rfind('/') + 1) + "libtxc_dxtn.so";
}

uint32_t le32(const uint8_t *p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }

}  // namespace

int main(int argc, char **argv)
{
	const char *infile = nullptr, *outfile = nullptr;
	std::string library = default_library();
	int opt;
	while ((opt = getopt(argc, argv, "i:o:l:")) != -1) {
		switch (opt) {
		case 'i': infile = optarg; break;
		case 'o': outfile = optarg; break;
		case 'l': library = optarg; break;
		default: return usage(argv[0]);
		}
	}
	void *lib = dlopen(library.c_str(), RTLD_NOW);
	if (!lib) {
		fprintf(stderr, "Cannot load library: %s\n", dlerror());
		return 1;
	}
	fetch_t *fetch1 = (fetch_t *)dlsym(lib, "fetch_2d_texel_rgba_dxt1");
	fetch_t *fetch3 = (fetch_t *)dlsym(lib, "fetch_2d_texel_rgba_dxt3");
	fetch_t *fetch5 = (fetch_t *)dlsym(lib, "fetch_2d_texel_rgba_dxt5");
	if (!dlsym(lib, "fetch_2d_texel_rgb_dxt1") || !fetch1 || !fetch3 || !fetch5) {
		fprintf(stderr, "The selected libtxc_dxtn.so does not contain all required symbols.");
		dlclose(lib);
		return 1;
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
	const int height = (int)le32(hdr + 12), width = (int)le32(hdr + 16);
	const uint32_t fourcc = le32(hdr + 84);
	fetch_t *fetch;
	unsigned fmt;
	int bs;
	if (fourcc == 0x31545844) fetch = fetch1, fmt = 0x83F1, bs = 8;
	else if (fourcc == 0x33545844) fetch = fetch3, fmt = 0x83F2, bs = 16;
	else if (fourcc == 0x35545844) fetch = fetch5, fmt = 0x83F3, bs = 16;
	else {
		fprintf(stderr, "Only DXT1, DXT3, DXT5 are supported!\n");
		return 1;
	}
	if (width <= 0 || height <= 0 || width > 65535 || height > 65535) {
		fprintf(stderr, "s2tc_decompress: bad image size %d x %d\n", width, height);
		return 1;
	}
	uint8_t t[18] = {0};
	t[2] = 2;
	t[12] = (uint8_t)(width % 256);
	t[13] = (uint8_t)(width / 256);
	t[14] = (uint8_t)(height % 256);
	t[15] = (uint8_t)(height / 256);
	t[16] = 32;
	t[17] = 0x28;
	fwrite(t, 18, 1, out);
	const size_t n = (size_t)((width + 3) / 4) * ((height + 3) / 4);
	std::vector<uint8_t> blocks(n * bs, 0);
	if (fread(blocks.data(), 1, blocks.size(), in) < blocks.size() && ferror(in)) return 2;
	std::vector<uint8_t> px((size_t)width * height * 4);
	decode_t *decode = (decode_t *)dlsym(lib, "s2tc_amd_decode_image");
	malloc_t *dmalloc = (malloc_t *)dlsym(lib, "s2tc_amd_malloc");
	free_t *dfree = (free_t *)dlsym(lib, "s2tc_amd_free");
	copy_t *h2d = (copy_t *)dlsym(lib, "s2tc_amd_memcpy_h2d");
	copy_t *d2h = (copy_t *)dlsym(lib, "s2tc_amd_memcpy_d2h");
	err_t *last_error = (err_t *)dlsym(lib, "s2tc_amd_last_error");
	if (decode && dmalloc && dfree && h2d && d2h && last_error) {
		void *db = nullptr, *dp = nullptr;
		int rc = dmalloc(&db, blocks.size());
		if (!rc) rc = dmalloc(&dp, px.size());
		if (!rc) rc = h2d(db, blocks.data(), blocks.size());
		if (!rc) rc = decode(db, fmt, width, height, dp, nullptr);
		if (!rc) rc = d2h(px.data(), dp, px.size());
		if (db) dfree(db);
		if (dp) dfree(dp);
		if (rc) {
			fprintf(stderr, "s2tc_decompress: GPU decode failed: %s\n", last_error());
			return 3;
		}
	} else {
		for (int y = 0; y < height; ++y)
			for (int x = 0; x < width; ++x) fetch(width, blocks.data(), x, y, &px[((size_t)y * width + x) * 4]);
	}
	for (size_t i = 0; i < (size_t)width * height; ++i) {
		const uint8_t r = px[4 * i];
		px[4 * i] = px[4 * i + 2];
		px[4 * i + 2] = r;
	}
	fwrite(px.data(), px.size(), 1, out);
	if (infile) fclose(in);
	if (outfile) fclose(out);
	return 0;
}
