// s2tc_compress.cpp - command line encoder: TGA in, DDS (with all mip levels) out.
//
// Same interface and output as the reference's s2tc_compress.c:
//   s2tc_compress [-i infile.tga] [-o outfile.dds] [-t DXT1|DXT3|DXT5] [-l libtxc_dxtn.so]
// stdin/stdout when -i/-o are missing; DXT1 by default (:574); the library
// is this project's libtxc_dxtn.so unless -l names another one, which is
// loaded with dlopen and must export tx_compress_dxtn (:39-55).
// The TGA loader accepts the reference's image types 1, 2, 3, 9, 10 and 11
// (:82-424), the DDS header is the reference's (:662-720), and every mip
// level is compressed with tx_compress_dxtn and then halved with the
// truncating box filter of Image_MipReduce32 (:427-493, :722-733).
#include <dlfcn.h>
#include <getopt.h>
#include <strings.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

typedef void (*tx_compress_dxtn_t)(int srccomps, int width, int height, const unsigned char *src, unsigned fmt,
                                   unsigned char *dest, int stride);

namespace {

int g_width = 0, g_height = 0;

// BGRA TGA decode (types 1, 2, 3, 9, 10, 11); returns an RGBA buffer or an
// empty vector after printing the reference's messages.
std::vector<uint8_t> load_tga(const std::vector<uint8_t> &file)
{
	std::vector<uint8_t> none;
	const size_t fs = file.size();
	if (fs < 19) return none;
	const uint8_t *f = file.data(), *end = f + fs;
	const int id_length = f[0], cmap_type = f[1], type = f[2];
	const int cmap_index = f[3] | f[4] << 8, cmap_length = f[5] | f[6] << 8, cmap_size = f[7];
	const int w = f[12] | f[13] << 8, h = f[14] | f[15] << 8;
	const int pixel_size = f[16], attributes = f[17];
	g_width = w;
	g_height = h;
	if (w > 32768 || h > 32768 || w <= 0 || h <= 0) {
		printf("LoadTGA: invalid size\n");
		return none;
	}
	const uint8_t *fin = f + 18 + id_length;
	uint32_t palette[256] = {0};
	if (cmap_type) {
		if (cmap_length > 256) {
			printf("LoadTGA: only up to 256 colormap_length supported\n");
			return none;
		}
		if (cmap_index) {
			printf("LoadTGA: colormap_index not supported\n");
			return none;
		}
		if (cmap_size == 24) {
			for (int x = 0; x < cmap_length; x++, fin += 3)
				palette[x] = fin[0] | fin[1] << 8 | fin[2] << 16 | 0xffu << 24;
		} else if (cmap_size == 32) {
			for (int x = 0; x < cmap_length; x++, fin += 4)
				palette[x] = fin[0] | fin[1] << 8 | fin[2] << 16 | (uint32_t)fin[3] << 24;
		} else {
			printf("LoadTGA: Only 32 and 24 bit colormap_size supported\n");
			return none;
		}
	}
	switch (type & ~8) {
	case 2:
		if (pixel_size != 24 && pixel_size != 32) {
			printf("LoadTGA: only 24bit and 32bit pixel sizes supported for type 2 and type 10 images\n");
			return none;
		}
		break;
	case 3:
		for (int x = 0; x < 256; x++) palette[x] = x | x << 8 | x << 16 | 0xffu << 24;
		/* fall through */
	case 1:
		if (pixel_size != 8) {
			printf("LoadTGA: only 8bit pixel size for type 1, 3, 9, and 11 images supported\n");
			return none;
		}
		break;
	default:
		printf("LoadTGA: Only type 1, 2, 3, 9, 10, and 11 targa RGB images supported, image_type = %i\n", type);
		return none;
	}
	if (attributes & 0x10) {
		printf("LoadTGA: origin must be in top left or bottom left, top right and bottom right are not supported\n");
		return none;
	}
	const int alphabits = attributes & 0x0f;
	if (alphabits != 8 && alphabits != 0) {
		printf("LoadTGA: only 0 or 8 attribute (alpha) bits supported\n");
		return none;
	}
	std::vector<uint32_t> img((size_t)w * h, 0);
	const bool bottom_up = (attributes & 0x20) == 0;
	auto row = [&](int y) { return img.data() + (size_t)(bottom_up ? h - 1 - y : y) * w; };
	const int pix_inc = (type & ~8) == 2 ? (pixel_size + 7) / 8 : 1;
	const bool keep_alpha = pixel_size == 32 && alphabits;
	auto bgra = [&](const uint8_t *p) -> uint32_t {
		return p[0] | p[1] << 8 | p[2] << 16 | (uint32_t)(keep_alpha ? p[3] : 255) << 24;
	};
	switch (type) {
	case 1:
	case 3:
		if (fin + (size_t)w * h * pix_inc > end) break;
		for (int y = 0; y < h; y++)
			for (int x = 0; x < w; x++) row(y)[x] = palette[*fin++];
		break;
	case 2:
		if (fin + (size_t)w * h * pix_inc > end) break;
		for (int y = 0; y < h; y++)
			for (int x = 0; x < w; x++, fin += pix_inc) row(y)[x] = bgra(fin);
		break;
	case 9:
	case 11:
	case 10:
		for (int y = 0; y < h; y++) {
			int x = 0;
			uint32_t *r = row(y);
			while (x < w) {
				if (fin >= end) break;
				int runlen = *fin++;
				if (runlen & 0x80) {
					runlen += 1 - 0x80;
					if (fin + pix_inc > end) break;
					if (x + runlen > w) break;
					const uint32_t v = type == 10 ? bgra(fin) : palette[*fin];
					fin += pix_inc;
					for (; runlen--; x++) r[x] = v;
				} else {
					runlen++;
					if (fin + (size_t)pix_inc * runlen > end) break;
					if (x + runlen > w) break;
					for (; runlen--; x++, fin += pix_inc) r[x] = type == 10 ? bgra(fin) : palette[*fin];
				}
			}
			if (x != w) {
				printf("LoadTGA: corrupt file\n");
				break;
			}
		}
		break;
	default:
		break;
	}
	// BGRA -> RGBA (s2tc_compress.c:635-639)
	std::vector<uint8_t> out((size_t)w * h * 4);
	for (size_t i = 0; i < img.size(); ++i) {
		const uint32_t v = img[i];
		out[4 * i + 0] = (uint8_t)(v >> 16);
		out[4 * i + 1] = (uint8_t)(v >> 8);
		out[4 * i + 2] = (uint8_t)v;
		out[4 * i + 3] = (uint8_t)(v >> 24);
	}
	return out;
}

// Image_MipReduce32 toward 1x1, in place
void mip_reduce(uint8_t *pic, int *width, int *height)
{
	const uint8_t *inrow = pic;
	uint8_t *out = pic;
	const int nextrow = *width * 4;
	if (*width > 1) {
		*width >>= 1;
		if (*height > 1) {
			*height >>= 1;
			for (int y = 0; y < *height; y++, inrow += nextrow * 2) {
				const uint8_t *in = inrow;
				for (int x = 0; x < *width; x++, out += 4, in += 8)
					for (int c = 0; c < 4; ++c)
						out[c] = (uint8_t)((in[c] + in[4 + c] + in[nextrow + c] + in[nextrow + 4 + c]) >> 2);
			}
		} else {
			for (int y = 0; y < *height; y++, inrow += nextrow) {
				const uint8_t *in = inrow;
				for (int x = 0; x < *width; x++, out += 4, in += 8)
					for (int c = 0; c < 4; ++c) out[c] = (uint8_t)((in[c] + in[4 + c]) >> 1);
			}
		}
	} else if (*height > 1) {
		*height >>= 1;
		for (int y = 0; y < *height; y++, inrow += nextrow * 2) {
			const uint8_t *in = inrow;
			for (int x = 0; x < *width; x++, out += 4, in += 4)
				for (int c = 0; c < 4; ++c) out[c] = (uint8_t)((in[c] + in[nextrow + c]) >> 1);
		}
	}
}

void put32(FILE *f, uint32_t v)
{
	uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
	fwrite(b, 4, 1, f);
}

int usage(const char *me)
{
	fprintf(stderr,
	        "usage:\n%s \n    [-i infile.tga]\n    [-o outfile.dds]\n    [-t {DXT1|DXT3|DXT5}]\n"
	        "    [-l path_to_libtxc_dxtn.so]\n",
	        me);
	return 1;
}

std::string default_library()
{
	// this project's libtxc_dxtn.so next to the executable
	char buf[4096];
	ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
	if (n <= 0) return "libtxc_dxtn.so";
	buf[n] = 0;
	std::string s(buf);
	return s.substr(0, s.rfind('/') + 1) + "libtxc_dxtn.so";
}

}  // namespace

int main(int argc, char **argv)
{
	const char *infile = nullptr, *outfile = nullptr;
	unsigned dxt = 0x83F1;  // GL_COMPRESSED_RGBA_S3TC_DXT1_EXT
	std::string library = default_library();
	int opt;
	while ((opt = getopt(argc, argv, "i:o:t:l:")) != -1) {
		switch (opt) {
		case 'i': infile = optarg; break;
		case 'o': outfile = optarg; break;
		case 't':
			if (!strcasecmp(optarg, "DXT1")) dxt = 0x83F1;
			else if (!strcasecmp(optarg, "DXT3")) dxt = 0x83F2;
			else if (!strcasecmp(optarg, "DXT5")) dxt = 0x83F3;
			else return usage(argv[0]);
			break;
		case 'l': library = optarg; break;
		default: return usage(argv[0]);
		}
	}
	void *lib = dlopen(library.c_str(), RTLD_NOW);
	if (!lib) {
		fprintf(stderr, "Cannot load library: %s\n", dlerror());
		return 1;
	}
	tx_compress_dxtn_t compress = (tx_compress_dxtn_t)dlsym(lib, "tx_compress_dxtn");
	if (!compress) {
		fprintf(stderr, "The selected libtxc_dxtn.so does not contain all required symbols.");
		dlclose(lib);
		return 1;
	}
	FILE *outfh = outfile ? fopen(outfile, "wb") : stdout;
	if (!outfh) {
		printf("opening output failed\n");
		return 2;
	}
	std::vector<uint8_t> file;
	{
		FILE *f = infile ? fopen(infile, "rb") : stdin;
		if (!f) {
			printf("FS_LoadFile failed\n");
			return 2;
		}
		uint8_t buf[65536];
		size_t n;
		while ((n = fread(buf, 1, sizeof buf, f)) > 0) file.insert(file.end(), buf, buf + n);
		if (infile) fclose(f);
	}
	std::vector<uint8_t> pic = load_tga(file);
	if (pic.empty()) {
		fprintf(stderr, "s2tc_compress: cannot decode the input image\n");
		return 2;
	}
	int mipcount = 0;
	while (g_width >= (1 << mipcount) || g_height >= (1 << mipcount)) ++mipcount;
	const int blocksize = dxt == 0x83F1 ? 8 : 16;
	const char *fourcc = dxt == 0x83F1 ? "DXT1" : dxt == 0x83F2 ? "DXT3" : "DXT5";
	bool alphapixels = false;
	for (size_t i = 0; i < (size_t)g_width * g_height; ++i)
		if (pic[4 * i + 3] != 255) {
			alphapixels = true;
			break;
		}
	// DDS header (s2tc_compress.c:662-720)
	fwrite("DDS ", 4, 1, outfh);
	put32(outfh, 0x7c);
	put32(outfh, 0x000a1007);
	put32(outfh, (uint32_t)g_height);
	put32(outfh, (uint32_t)g_width);
	put32(outfh, (uint32_t)(((g_width + 3) / 4) * ((g_height + 3) / 4) * blocksize));
	put32(outfh, 0);
	put32(outfh, (uint32_t)mipcount);
	for (int i = 0; i < 11; ++i) put32(outfh, 0);
	put32(outfh, 0x20);
	put32(outfh, alphapixels ? 5 : 4);
	fwrite(fourcc, 4, 1, outfh);
	for (int i = 0; i < 5; ++i) put32(outfh, 0);
	put32(outfh, 0x00401008);
	for (int i = 0; i < 4; ++i) put32(outfh, 0);
	int w = g_width, h = g_height;
	std::vector<uint8_t> obuf;
	for (;;) {
		const int bw = (w + 3) / 4, bh = (h + 3) / 4;
		obuf.assign((size_t)bw * bh * blocksize, 0);
		compress(4, w, h, pic.data(), dxt, obuf.data(), bw * blocksize);
		fwrite(obuf.data(), obuf.size(), 1, outfh);
		if (w == 1 && h == 1) break;
		mip_reduce(pic.data(), &w, &h);
	}
	if (outfile) fclose(outfh);
	return 0;
}
