// synth.cpp - deterministic synthetic inputs for the benchmark workloads
// (libs2tc_synth.so; not part of the encoder library).
//
// Everything is integer arithmetic or IEEE double arithmetic in a fixed
// order with this file's own sin/exp polynomials (no libm, no FMA
// contraction: built with -ffp-contract=off), so the same seed gives the
// same bytes on every x86-64 host.
//
// Workloads:
//   sinus     128 + 100*sin(x/17 + phase_c)*cos(y/23) + U[-24,24] per channel,
//             alpha 255 (the 256x256 single-image case)
//   tiles     64x64 tiles, each a linear gradient between two random colours
//             plus U[-8,8]; alpha = radial falloff per tile with 10% of the
//             texels forced to 0 or 255 (the 4096x4096 DXT5 case)
//   normalmap normals of a height field made of random Gaussians
//             (sigma 8..128 px scaled to the image, amplitude +-1), central
//             differences, round(n*127.5 + 127.5) (the 8192x8192 case)
//   batch     texture specs: sides 2^k with k uniform in 2..11 per side and
//             |kx - ky| <= 3 (rejection), content kind uniform over the three
//             above, DXT1/DXT3/DXT5 with probability 0.3/0.3/0.4, distance
//             uniform over the 8 modes; for the serving mix 70% quick mode
//             (nrandom -1, ALWAYS) and 30% nrandom 64 with LOOP.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define SYN_API extern "C" __attribute__((visibility("default")))

namespace {

struct Rng {  // splitmix64
	uint64_t s;
	explicit Rng(uint64_t seed) : s(seed) {}
	uint64_t next()
	{
		uint64_t z = (s += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	}
	int uniform(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }  // [lo, hi]
	double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }    // [0, 1)
};

uint64_t mix64(uint64_t a, uint64_t b)
{
	Rng r(a ^ (b * 0xD1B54A32D192ED03ull));
	r.next();
	return r.next();
}

const double kPi = 3.14159265358979323846;

// sin on [-pi, pi] after reduction by 2*pi; degree-15 odd Taylor polynomial
double dsin(double x)
{
	const double twopi = 2.0 * kPi;
	double k = std::floor(x / twopi + 0.5);
	x = x - k * twopi;
	const double x2 = x * x;
	double p = -7.647163731819816e-13;
	p = p * x2 + 1.6059043836821613e-10;
	p = p * x2 - 2.505210838544172e-08;
	p = p * x2 + 2.7557319223985893e-06;
	p = p * x2 - 0.0001984126984126984;
	p = p * x2 + 0.008333333333333333;
	p = p * x2 - 0.16666666666666666;
	p = p * x2 + 1.0;
	return p * x;
}
double dcos(double x) { return dsin(x + 0.5 * kPi); }

// exp(-t) for t >= 0: 2^-k * exp(-r) with r in [0, ln 2)
double dexpneg(double t)
{
	if (t > 700.0) return 0.0;
	const double ln2 = 0.6931471805599453;
	double k = std::floor(t / ln2);
	double r = t - k * ln2;
	double p = 1.0, term = 1.0;
	for (int i = 1; i < 18; ++i) {
		term = term * (-r) / (double)i;
		p = p + term;
	}
	double s = 1.0;
	int ki = (int)k;
	while (ki >= 60) {
		s *= 8.673617379884035e-19;  // 2^-60
		ki -= 60;
	}
	s *= 1.0 / (double)(1ull << ki);
	return p * s;
}

inline uint8_t clamp255(int v) { return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v); }

void parallel_rows(int h, int nthreads, const std::function<void(int, int)> &fn)
{
	if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
	nthreads = std::max(1, std::min(nthreads, 64));
	if (nthreads == 1 || h < 64) {
		fn(0, h);
		return;
	}
	std::atomic<int> next{0};
	const int chunk = 16;
	std::vector<std::thread> th;
	for (int t = 0; t < nthreads; ++t)
		th.emplace_back([&] {
			for (;;) {
				int y0 = next.fetch_add(chunk);
				if (y0 >= h) break;
				fn(y0, std::min(h, y0 + chunk));
			}
		});
	for (auto &t : th) t.join();
}
}  // namespace

SYN_API void s2tc_synth_sinus(uint8_t *rgba, int w, int h, uint64_t seed, int nthreads)
{
	const double phase = (double)(seed % 1000) * 0.001 * 2.0 * kPi;
	parallel_rows(h, nthreads, [&](int y0, int y1) {
		for (int y = y0; y < y1; ++y) {
			Rng r(mix64(seed, (uint64_t)y));
			const double cy = dcos((double)y / 23.0);
			for (int x = 0; x < w; ++x) {
				uint8_t *p = rgba + ((size_t)y * w + x) * 4;
				for (int c = 0; c < 3; ++c) {
					const double v = 128.0 + 100.0 * dsin((double)x / 17.0 + phase + 2.0943951023931953 * c) * cy;
					p[c] = clamp255((int)std::floor(v) + r.uniform(-24, 24));
				}
				p[3] = 255;
			}
		}
	});
}

SYN_API void s2tc_synth_tiles(uint8_t *rgba, int w, int h, uint64_t seed, int nthreads)
{
	const int T = 64;
	const int tw = (w + T - 1) / T;
	parallel_rows(h, nthreads, [&](int y0, int y1) {
		for (int y = y0; y < y1; ++y) {
			Rng noise(mix64(seed ^ 0x5151, (uint64_t)y));
			for (int x = 0; x < w; ++x) {
				const int tx = x / T, ty = y / T;
				Rng tr(mix64(seed, (uint64_t)(ty * tw + tx)));
				int ca[3], cb[3];
				for (int c = 0; c < 3; ++c) ca[c] = tr.uniform(0, 255);
				for (int c = 0; c < 3; ++c) cb[c] = tr.uniform(0, 255);
				const double ang = tr.unit() * 2.0 * kPi;
				const double dx = dcos(ang), dy = dsin(ang);
				const double u = ((double)(x % T) - 31.5) / 63.0, v = ((double)(y % T) - 31.5) / 63.0;
				double t = 0.5 + (u * dx + v * dy);
				t = t < 0 ? 0 : t > 1 ? 1 : t;
				uint8_t *p = rgba + ((size_t)y * w + x) * 4;
				for (int c = 0; c < 3; ++c)
					p[c] = clamp255((int)std::floor(ca[c] + (cb[c] - ca[c]) * t + 0.5) + noise.uniform(-8, 8));
				const double rr = std::sqrt(u * u + v * v) / 0.7071067811865476;  // 0 at centre, 1 at corner
				int a = (int)std::floor(255.0 * (1.0 - (rr > 1 ? 1 : rr)) + 0.5);
				const uint64_t sp = noise.next();
				if (sp % 10 == 0) a = (sp >> 8) & 1 ? 255 : 0;
				p[3] = clamp255(a);
			}
		}
	});
}

SYN_API void s2tc_synth_normalmap(uint8_t *out, int comps, int w, int h, int ngauss, uint64_t seed, int nthreads)
{
	// Gaussians: centre, sigma (8..128 px at 8192, scaled with the image), amplitude +-1
	struct G {
		int cx, cy, r0x, r1x, r0y, r1y;
		double amp;
		std::vector<float> wx, wy;
	};
	std::vector<G> gs((size_t)ngauss);
	Rng r(seed);
	const double scale = (double)std::max(w, h) / 8192.0;
	for (auto &g : gs) {
		g.cx = r.uniform(0, w - 1);
		g.cy = r.uniform(0, h - 1);
		const double sigma = std::max(1.0, (8.0 + 120.0 * r.unit()) * std::max(scale, 1.0 / 64.0));
		g.amp = r.unit() * 2.0 - 1.0;
		const int rad = (int)std::ceil(4.0 * sigma);
		g.r0x = std::max(0, g.cx - rad);
		g.r1x = std::min(w - 1, g.cx + rad);
		g.r0y = std::max(0, g.cy - rad);
		g.r1y = std::min(h - 1, g.cy + rad);
		for (int x = g.r0x; x <= g.r1x; ++x) {
			const double d = (double)(x - g.cx);
			g.wx.push_back((float)(g.amp * dexpneg(d * d / (2.0 * sigma * sigma))));
		}
		for (int y = g.r0y; y <= g.r1y; ++y) {
			const double d = (double)(y - g.cy);
			g.wy.push_back((float)dexpneg(d * d / (2.0 * sigma * sigma)));
		}
	}
	std::vector<float> hf((size_t)w * h, 0.0f);
	// each row sums the Gaussians in index order: the result does not depend
	// on the thread split
	parallel_rows(h, nthreads, [&](int y0, int y1) {
		for (int y = y0; y < y1; ++y) {
			float *row = hf.data() + (size_t)y * w;
			for (const G &g : gs) {
				if (y < g.r0y || y > g.r1y) continue;
				const float wy = g.wy[(size_t)(y - g.r0y)];
				for (int x = g.r0x; x <= g.r1x; ++x) row[x] += g.wx[(size_t)(x - g.r0x)] * wy;
			}
		}
	});
	const double k = 64.0;
	parallel_rows(h, nthreads, [&](int y0, int y1) {
		for (int y = y0; y < y1; ++y) {
			for (int x = 0; x < w; ++x) {
				const int xl = std::max(0, x - 1), xr = std::min(w - 1, x + 1);
				const int yu = std::max(0, y - 1), yd = std::min(h - 1, y + 1);
				const double gx = ((double)hf[(size_t)y * w + xr] - hf[(size_t)y * w + xl]) / std::max(1, xr - xl);
				const double gy = ((double)hf[(size_t)yd * w + x] - hf[(size_t)yu * w + x]) / std::max(1, yd - yu);
				double nx = -gx * k, ny = -gy * k, nz = 1.0;
				const double inv = 1.0 / std::sqrt(nx * nx + ny * ny + nz * nz);
				nx *= inv;
				ny *= inv;
				nz *= inv;
				uint8_t *p = out + ((size_t)y * w + x) * comps;
				p[0] = clamp255((int)std::floor(nx * 127.5 + 127.5 + 0.5));
				p[1] = clamp255((int)std::floor(ny * 127.5 + 127.5 + 0.5));
				p[2] = clamp255((int)std::floor(nz * 127.5 + 127.5 + 0.5));
				if (comps == 4) p[3] = 255;
			}
		}
	});
}

struct S2tcTexSpec {
	int width, height;
	int dxt, cd, nrandom, refine;
	int kind;  // 0 sinus, 1 tiles, 2 normalmap
	int pad;
	uint64_t seed;
};

// serve_mix = 0: batch of mixed sizes/formats/distances, quick settings;
// serve_mix = 1: the same plus 30% nrandom 64 + LOOP.
SYN_API void s2tc_synth_batch_specs(S2tcTexSpec *specs, int n, uint64_t seed, int serve_mix)
{
	Rng r(seed);
	for (int i = 0; i < n; ++i) {
		S2tcTexSpec &s = specs[i];
		int kx, ky;
		do {
			kx = r.uniform(2, 11);
			ky = r.uniform(2, 11);
		} while (kx - ky > 3 || ky - kx > 3);
		s.width = 1 << kx;
		s.height = 1 << ky;
		s.kind = r.uniform(0, 2);
		const int f = r.uniform(0, 9);
		s.dxt = f < 3 ? 0 : f < 6 ? 1 : 2;
		s.cd = r.uniform(0, 7);
		s.nrandom = -1;
		s.refine = 1;
		const int heavy = r.uniform(0, 9);
		if (serve_mix && heavy < 3) {
			s.nrandom = 64;
			s.refine = 2;
		}
		s.pad = 0;
		s.seed = r.next();
	}
}

SYN_API void s2tc_synth_texture(uint8_t *rgba, const S2tcTexSpec *s, int nthreads)
{
	switch (s->kind) {
	case 0: s2tc_synth_sinus(rgba, s->width, s->height, s->seed, nthreads); break;
	case 1: s2tc_synth_tiles(rgba, s->width, s->height, s->seed, nthreads); break;
	default: {
		const int ng = std::max(4, (int)((double)512 * (double)s->width * s->height / (8192.0 * 8192.0)));
		s2tc_synth_normalmap(rgba, 4, s->width, s->height, ng, s->seed, nthreads);
	}
	}
}

// Image_MipReduce32 toward 1x1 (s2tc_compress.c:427-493): truncating box
// filter, odd rows/columns dropped; in may equal out.
SYN_API void s2tc_synth_mip_reduce(const uint8_t *in, uint8_t *out, int *width, int *height)
{
	const uint8_t *inrow = in;
	const int nextrow = *width * 4;
	if (*width > 1) {
		*width >>= 1;
		if (*height > 1) {
			*height >>= 1;
			for (int y = 0; y < *height; y++, inrow += nextrow * 2) {
				const uint8_t *p = inrow;
				for (int x = 0; x < *width; x++, out += 4, p += 8)
					for (int c = 0; c < 4; ++c)
						out[c] = (uint8_t)((p[c] + p[4 + c] + p[nextrow + c] + p[nextrow + 4 + c]) >> 2);
			}
		} else {
			for (int y = 0; y < *height; y++, inrow += nextrow) {
				const uint8_t *p = inrow;
				for (int x = 0; x < *width; x++, out += 4, p += 8)
					for (int c = 0; c < 4; ++c) out[c] = (uint8_t)((p[c] + p[4 + c]) >> 1);
			}
		}
	} else if (*height > 1) {
		*height >>= 1;
		for (int y = 0; y < *height; y++, inrow += nextrow * 2) {
			const uint8_t *p = inrow;
			for (int x = 0; x < *width; x++, out += 4, p += 4)
				for (int c = 0; c < 4; ++c) out[c] = (uint8_t)((p[c] + p[nextrow + c]) >> 1);
		}
	}
}

// Number of mip levels down to 1x1: floor(log2(max(w, h))) + 1 (s2tc_compress.c:640-642)
SYN_API int s2tc_synth_mip_count(int w, int h)
{
	int n = 0;
	while (w >= (1 << n) || h >= (1 << n)) ++n;
	return n;
}
