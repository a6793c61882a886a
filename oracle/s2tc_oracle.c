/*
 * s2tc_oracle.c - CPU restatement of the reference S2TC encoder (checker).
 *
 * TEST INFRASTRUCTURE ONLY: see s2tc_oracle.h.  Built with
 * -O2 -ffp-contract=off so float expressions round exactly as the
 * reference's (x86-64 SSE, no FMA).  File:line citations refer to the
 * reference sources (s2tc_algorithm.cpp, s2tc_libtxc_dxtn.cpp,
 * s2tc_compress.c, s2tc_common.h).
 *
 * Signed overflow in the reference (SRGB distance, s2tc_algorithm.cpp:277-280,
 * and the int sums of reduce_colors_inplace) wraps in practice; here every
 * such operation goes through uint32 so the wrap is defined behaviour.
 */
#include "s2tc_oracle.h"
#include "../include/s2tc_prng.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int r, g, b; } ocol;   /* color_t: signed char r,g,b (s2tc_algorithm.cpp:50-53) */

static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
/* SHRR(a,n) = ((a) + (1 << ((n)-1))) >> (n), arithmetic shift (s2tc_algorithm.cpp:215) */
static inline int32_t shrr(int32_t a, int n) { return wadd(a, 1 << (n - 1)) >> n; }
static inline signed char sc(int v) { return (signed char)v; }

/* ---- distances: s2tc_algorithm.cpp:217-361 ---- */

static int dist_avg(ocol a, ocol b)   /* :217-223 */
{
	int dr = a.r - b.r, dg = a.g - b.g, db = a.b - b.b;
	return ((dr * dr) << 2) + dg * dg + ((db * db) << 2);
}
static int dist_w0avg(ocol a, ocol b) /* :225-232 */
{
	int dr = a.r - b.r, dg = a.g - b.g, db = a.b - b.b;
	return dr * dr + dg * dg + db * db;
}
static int dist_wavg(ocol a, ocol b)  /* :234-241 */
{
	int dr = a.r - b.r, dg = a.g - b.g, db = a.b - b.b;
	return ((dr * dr) << 2) + ((dg * dg) << 2) + (db * db);
}
static int dist_yuvlike(ocol a, ocol b, int kr, int kg, int kb) /* :243-267 */
{
	int32_t dr = a.r - b.r, dg = a.g - b.g, db = a.b - b.b;
	int32_t y = wadd(wadd(wmul(dr, kr), wmul(dg, kg)), wmul(db, kb));
	int32_t u = wsub(wmul(dr, 202), y);
	int32_t v = wsub(wmul(db, 202), y);
	return wadd(wadd(wmul(wmul(y, y), 2), shrr(wmul(u, u), 3)), shrr(wmul(v, v), 4));
}
static int dist_srgb(ocol a, ocol b)  /* :269-283, wraps int32 */
{
	int32_t dr = a.r * a.r - b.r * b.r;
	int32_t dg = a.g * a.g - b.g * b.g;
	int32_t db = a.b * a.b - b.b * b.b;
	int32_t y = wadd(wadd(wmul(dr, 84), wmul(dg, 72)), wmul(db, 28));
	int32_t u = wsub(wmul(dr, 409), y);
	int32_t v = wsub(wmul(db, 409), y);
	int32_t sy = wmul(shrr(y, 3), shrr(y, 4));
	int32_t su = wmul(shrr(u, 3), shrr(u, 4));
	int32_t sv = wmul(shrr(v, 3), shrr(v, 4));
	return wadd(wadd(shrr(sy, 4), shrr(su, 8)), shrr(sv, 9));
}
int oracle_srgb_get_y(int r, int g, int b) /* :285-296 */
{
	int y = 37 * (r * r * 84 + g * g * 72 + b * b * 28);
	return (int)(sqrtf((float)y) + 0.5f);
}
static int dist_srgb_mixed(ocol a, ocol b) /* :298-315 */
{
	int ay = oracle_srgb_get_y(a.r, a.g, a.b);
	int by = oracle_srgb_get_y(b.r, b.g, b.b);
	int au = a.r * 191 - ay, av = a.b * 191 - ay;
	int bu = b.r * 191 - by, bv = b.b * 191 - by;
	int32_t y = ay - by, u = au - bu, v = av - bv;
	return wadd(wadd(wmul(wmul(y, y), 8), shrr(wmul(u, u), 1)), shrr(wmul(v, v), 2));
}
static int dist_normalmap(ocol a, ocol b) /* :317-354 */
{
	float ca[3], cb[3], n;
	ca[0] = (float)a.r / 31.0f * 2.0f - 1.0f;
	ca[1] = (float)a.g / 63.0f * 2.0f - 1.0f;
	ca[2] = (float)a.b / 31.0f * 2.0f - 1.0f;
	cb[0] = (float)b.r / 31.0f * 2.0f - 1.0f;
	cb[1] = (float)b.g / 63.0f * 2.0f - 1.0f;
	cb[2] = (float)b.b / 31.0f * 2.0f - 1.0f;
	n = ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2];
	if (n > 0) {
		n = 1.0f / sqrtf(n);
		ca[0] *= n; ca[1] *= n; ca[2] *= n;
	}
	n = cb[0] * cb[0] + cb[1] * cb[1] + cb[2] * cb[2];
	if (n > 0) {
		n = 1.0f / sqrtf(n);
		cb[0] *= n; cb[1] *= n; cb[2] *= n;
	}
	return (int)(100000.0f * ((cb[0] - ca[0]) * (cb[0] - ca[0])
	                          + (cb[1] - ca[1]) * (cb[1] - ca[1])
	                          + (cb[2] - ca[2]) * (cb[2] - ca[2])));
}

static int cdist(int cd, ocol a, ocol b)
{
	switch (cd) {
	case O_RGB: return dist_yuvlike(a, b, 42, 72, 14);
	case O_YUV: return dist_yuvlike(a, b, 60, 59, 22);
	case O_SRGB: return dist_srgb(a, b);
	case O_SRGB_MIXED: return dist_srgb_mixed(a, b);
	case O_AVG: return dist_avg(a, b);
	case O_W0AVG: return dist_w0avg(a, b);
	case O_NORMALMAP: return dist_normalmap(a, b);
	case O_WAVG:
	default: return dist_wavg(a, b);
	}
}

int oracle_color_dist(int cd, int ar, int ag, int ab, int br, int bg, int bb)
{
	ocol a = { sc(ar), sc(ag), sc(ab) }, b = { sc(br), sc(bg), sc(bb) };
	return cdist(cd, a, b);
}

static inline int adist(int a, int b) { return (a - b) * (a - b); } /* alpha_dist :358-361 */

/* ---- color_t order and lattice steps: :66-131 ---- */
static inline int ceq(ocol a, ocol b) { return a.r == b.r && a.g == b.g && a.b == b.b; }
static inline int clt(ocol a, ocol b)
{
	signed char d = sc(a.r - b.r);
	if (d) return d < 0;
	d = sc(a.g - b.g);
	if (d) return d < 0;
	d = sc(a.b - b.b);
	return d < 0;
}
static void cdec(ocol *c) /* operator-- :82-106 */
{
	if (c->b > 0) --c->b;
	else if (c->g > 0) { c->b = 31; --c->g; }
	else if (c->r > 0) { c->b = 31; c->g = 63; --c->r; }
	else { c->b = 31; c->g = 63; c->r = 31; }
}
static void cinc(ocol *c) /* operator++ :107-131 */
{
	if (c->b < 31) ++c->b;
	else if (c->g < 63) { c->b = 0; ++c->g; }
	else if (c->r < 31) { c->b = 0; c->g = 0; ++c->r; }
	else { c->b = 0; c->g = 0; c->r = 0; }
}
static const ocol CMAX = { 31, 63, 31 }; /* :137-138 */

/* ---- reduce_colors_inplace :363-414 (with the "bestsum < 0" update test) ---- */
static void reduce_colors(ocol *c, int n, int m, int cd)
{
	int i, j, k, bestsum = -1, besti = 0, bestj = 1;
	int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)m * (size_t)n);
	for (i = 0; i < n; ++i) {
		d[i * n + i] = 0;
		for (j = i + 1; j < n; ++j) {
			int v = cdist(cd, c[i], c[j]);
			d[i * n + j] = v;
			d[j * n + i] = v;
		}
	}
	for (; i < m; ++i)
		for (j = 0; j < n; ++j)
			d[i * n + j] = cdist(cd, c[i], c[j]);
	for (i = 0; i < m; ++i)
		for (j = i + 1; j < m; ++j) {
			int32_t sum = 0;
			for (k = 0; k < n; ++k) {
				int32_t di = d[i * n + k], dj = d[j * n + k];
				sum = wadd(sum, dj < di ? dj : di);
			}
			if (bestsum < 0 || sum < bestsum) {
				bestsum = sum;
				besti = i;
				bestj = j;
			}
		}
	{
		ocol c0 = c[besti];
		c[1] = c[bestj];
		c[0] = c0;
	}
	free(d);
}

/* ---- reduce_colors_inplace_2fixpoints :415-478, alpha with fixed 0 and 255 ---- */
static void reduce_alpha_2fix(uint8_t *c, int n, int m)
{
	int i, j, k, bestsum = -1, besti = 0, bestj = 1;
	int32_t *d = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m + 2) * (size_t)n);
	for (i = 0; i < n; ++i) {
		d[i * n + i] = 0;
		for (j = i + 1; j < n; ++j) {
			int v = adist(c[i], c[j]);
			d[i * n + j] = v;
			d[j * n + i] = v;
		}
	}
	for (; i < m; ++i)
		for (j = 0; j < n; ++j)
			d[i * n + j] = adist(c[i], c[j]);
	for (j = 0; j < n; ++j) d[m * n + j] = adist(0, c[j]);
	for (j = 0; j < n; ++j) d[(m + 1) * n + j] = adist(255, c[j]);
	for (i = 0; i < m; ++i)
		for (j = i + 1; j < m; ++j) {
			int32_t sum = 0;
			for (k = 0; k < n; ++k) {
				int32_t di = d[i * n + k], dj = d[j * n + k];
				int32_t d0 = d[m * n + k], d1 = d[(m + 1) * n + k];
				int32_t a = dj < di ? dj : di;
				int32_t b = d1 < d0 ? d1 : d0;
				sum = wadd(sum, b < a ? b : a);
			}
			if (bestsum < 0 || sum < bestsum) {
				bestsum = sum;
				besti = i;
				bestj = j;
			}
		}
	if (besti != 0) c[0] = c[besti];
	if (bestj != 1) c[1] = c[bestj];
	free(d);
}

/* ---- refinement accumulator s2tc_evaluate_colors_result_t<...,1> :515-552 ---- */
typedef struct { int n0, n1; int s0[3], s1[3]; } acc_t;

static void acc_add(acc_t *r, int l, const int *v, int nc)
{
	int q;
	if (l) { ++r->n1; for (q = 0; q < nc; ++q) r->s1[q] += v[q]; }
	else { ++r->n0; for (q = 0; q < nc; ++q) r->s0[q] += v[q]; }
}
/* ((S << 1) + n) / (n << 1): round-half-up mean; colour components are then
 * masked &31/&63/&31 by bigcolor_t::operator color_t (:201-208). */
static int acc_eval_color(const acc_t *r, ocol *a, ocol *b)
{
	if (!r->n0 && !r->n1) return 0;
	if (r->n0) {
		a->r = sc((((r->s0[0] << 1) + r->n0) / (r->n0 << 1)) & 31);
		a->g = sc((((r->s0[1] << 1) + r->n0) / (r->n0 << 1)) & 63);
		a->b = sc((((r->s0[2] << 1) + r->n0) / (r->n0 << 1)) & 31);
	}
	if (r->n1) {
		b->r = sc((((r->s1[0] << 1) + r->n1) / (r->n1 << 1)) & 31);
		b->g = sc((((r->s1[1] << 1) + r->n1) / (r->n1 << 1)) & 63);
		b->b = sc((((r->s1[2] << 1) + r->n1) / (r->n1 << 1)) & 31);
	}
	return 1;
}
static int acc_eval_alpha(const acc_t *r, uint8_t *a, uint8_t *b)
{
	if (!r->n0 && !r->n1) return 0;
	if (r->n0) *a = (uint8_t)(((r->s0[0] << 1) + r->n0) / (r->n0 << 1));
	if (r->n1) *b = (uint8_t)(((r->s1[0] << 1) + r->n1) / (r->n1 << 1));
	return 1;
}

/* ---- s2tc_try_encode_block :582-643, colour variant (bpp 2, no 0/255) ---- */
static uint32_t try_color(uint32_t *out, acc_t *res, int cd, int have_trans,
                          const uint8_t *in, int iw, int w, int h, ocol r0, ocol r1)
{
	uint32_t score = 0;
	int x, y;
	for (x = 0; x < w; ++x) for (y = 0; y < h; ++y) {
		int i = y * 4 + x;
		const uint8_t *pix = &in[(y * iw + x) * 4];
		ocol c;
		int best = 0, bestdist, dd;
		if (have_trans && pix[3] == 0) {
			*out |= 3u << (2 * i);
			continue;
		}
		c.r = sc(pix[0]); c.g = sc(pix[1]); c.b = sc(pix[2]);
		bestdist = cdist(cd, c, r0);
		dd = cdist(cd, c, r1);
		if (dd < bestdist) { bestdist = dd; best = 1; }
		if (res) { int v[3] = { c.r, c.g, c.b }; acc_add(res, best, v, 3); }
		*out |= (uint32_t)best << (2 * i);
		score += (uint32_t)bestdist;
	}
	return score;
}

/* ---- s2tc_try_encode_block, alpha variant (bpp 3, have_0_255) ---- */
static uint32_t try_alpha(uint64_t *out, acc_t *res, const uint8_t *in, int iw, int w, int h,
                          uint8_t r0, uint8_t r1)
{
	uint32_t score = 0;
	int x, y;
	for (x = 0; x < w; ++x) for (y = 0; y < h; ++y) {
		int i = y * 4 + x;
		int a = in[(y * iw + x) * 4 + 3];
		int best = 0, bestdist = adist(a, r0), dd = adist(a, r1);
		if (dd < bestdist) { bestdist = dd; best = 1; }
		dd = adist(a, 0);
		if (dd <= bestdist) {
			*out |= (uint64_t)6 << (3 * i);
			score += (uint32_t)dd;
			continue;
		}
		dd = adist(a, 255);
		if (dd <= bestdist) {
			*out |= (uint64_t)7 << (3 * i);
			score += (uint32_t)dd;
			continue;
		}
		if (res) { int v[1] = { a }; acc_add(res, best, v, 1); }
		*out |= (uint64_t)best << (3 * i);
		score += (uint32_t)bestdist;
	}
	return score;
}

static inline int get2(uint32_t bits, int i) { return (int)((bits >> (2 * i)) & 3); }
static inline void set2(uint32_t *bits, int i, int v) { *bits = (*bits & ~(3u << (2 * i))) | ((uint32_t)v << (2 * i)); }
static inline int get3(uint64_t bits, int i) { return (int)((bits >> (3 * i)) & 7); }
static inline void set3(uint64_t *bits, int i, int v) { *bits = (*bits & ~((uint64_t)7 << (3 * i))) | ((uint64_t)v << (3 * i)); }

/* equal-endpoint fix + ordering of the colour refiners :796-813 / :828-845 */
static void color_fix_order(uint32_t *out, ocol *c0, ocol *c1, int have_trans)
{
	int i;
	if (ceq(*c0, *c1)) {
		if (ceq(*c0, CMAX)) cdec(c1); else cinc(c1);
		for (i = 0; i < 16; ++i)
			if (!(get2(*out, i) == 1)) set2(out, i, 0);
	}
	if (have_trans ? clt(*c1, *c0) : clt(*c0, *c1)) {
		ocol t = *c0; *c0 = *c1; *c1 = t;
		for (i = 0; i < 16; ++i)
			if (!(get2(*out, i) & 2)) *out ^= 1u << (2 * i);
	}
}

/* colour refine NEVER/ALWAYS/LOOP :767-860 */
static uint32_t color_refine(int refine, int cd, int have_trans, const uint8_t *in, int iw,
                             int w, int h, ocol *c0, ocol *c1)
{
	uint32_t out = 0;
	if (refine == O_REFINE_NEVER) {
		if (have_trans ? clt(*c1, *c0) : clt(*c0, *c1)) { ocol t = *c0; *c0 = *c1; *c1 = t; }
		try_color(&out, NULL, cd, have_trans, in, iw, w, h, *c0, *c1);
		return out;
	}
	if (refine == O_REFINE_LOOP) {
		uint32_t out2 = 0, s = 0x7FFFFFFF;
		ocol n0 = *c0, n1 = *c1;
		for (;;) {
			acc_t r2; uint32_t s2;
			memset(&r2, 0, sizeof r2);
			s2 = try_color(&out2, &r2, cd, have_trans, in, iw, w, h, n0, n1);
			if (s2 < s) {
				out = out2; s = s2; *c0 = n0; *c1 = n1;
				if (!acc_eval_color(&r2, &n0, &n1)) break;
			} else break;
			out2 = 0;
		}
	} else { /* ALWAYS */
		acc_t r2;
		memset(&r2, 0, sizeof r2);
		try_color(&out, &r2, cd, have_trans, in, iw, w, h, *c0, *c1);
		acc_eval_color(&r2, c0, c1);
	}
	color_fix_order(&out, c0, c1, have_trans);
	return out;
}

/* alpha equal fix + ordering :673-705 / :719-751 */
static void alpha_fix_order(uint64_t *out, uint8_t *a0, uint8_t *a1)
{
	int i;
	if (*a1 == *a0) {
		if (*a0 == 255) --*a1; else ++*a1;
		for (i = 0; i < 16; ++i)
			if (get3(*out, i) == 1) set3(out, i, 0);
	}
	if (*a1 < *a0) {
		uint8_t t = *a0; *a0 = *a1; *a1 = t;
		for (i = 0; i < 16; ++i) {
			int v = get3(*out, i);
			if (v == 0) set3(out, i, 1);
			else if (v == 1) set3(out, i, 0);
			else if (v == 6 || v == 7) { }
			else set3(out, i, 7 - v);
		}
	}
}

/* alpha refine NEVER/ALWAYS/LOOP :645-765 */
static uint64_t alpha_refine(int refine, const uint8_t *in, int iw, int w, int h, uint8_t *a0, uint8_t *a1)
{
	uint64_t out = 0;
	if (refine == O_REFINE_NEVER) {
		if (*a1 < *a0) { uint8_t t = *a0; *a0 = *a1; *a1 = t; }
		try_alpha(&out, NULL, in, iw, w, h, *a0, *a1);
		return out;
	}
	if (refine == O_REFINE_LOOP) {
		uint64_t out2 = 0;
		uint32_t s = 0x7FFFFFFF;
		uint8_t n0 = *a0, n1 = *a1;
		for (;;) {
			acc_t r2; uint32_t s2;
			memset(&r2, 0, sizeof r2);
			s2 = try_alpha(&out2, &r2, in, iw, w, h, n0, n1);
			if (s2 < s) {
				out = out2; s = s2; *a0 = n0; *a1 = n1;
				if (!acc_eval_alpha(&r2, &n0, &n1)) break;
			} else break;
			out2 = 0;
		}
	} else {
		acc_t r2;
		memset(&r2, 0, sizeof r2);
		try_alpha(&out, &r2, in, iw, w, h, *a0, *a1);
		acc_eval_alpha(&r2, a0, a1);
	}
	alpha_fix_order(&out, a0, a1);
	return out;
}

static void pack565(uint8_t *o, ocol c) /* :1046-1049 */
{
	o[0] = (uint8_t)(((c.g & 0x07) << 5) | c.b);
	o[1] = (uint8_t)((c.r << 3) | (c.g >> 3));
}

/* ---- s2tc_encode_block<dxt, ColorDist, mode, refine> :872-1108 with the
 *      dispatch rules of :1110-1194 ---- */
void oracle_encode_block(uint8_t *out, const uint8_t *rgba, int iw, int w, int h,
                         int dxt, int cd, int nrandom, int refine, uint32_t bkey)
{
	int cap = 16 + (nrandom >= 0 ? nrandom : 0);
	ocol *c = (ocol *)calloc((size_t)cap + 2, sizeof(ocol));
	uint8_t *ca = (uint8_t *)calloc((size_t)cap + 2, 1);   /* zero: pins ca[1] (see header) */
	int x, y, fast;
	uint32_t cbits;

	/* dispatch defaults :1120-1122, :1156-1158, :1183-1186 */
	if (refine != O_REFINE_NEVER && refine != O_REFINE_LOOP) refine = O_REFINE_ALWAYS;
	if (dxt != O_DXT1 && dxt != O_DXT3) dxt = O_DXT5;
	if (cd < O_RGB || cd > O_NORMALMAP) cd = O_WAVG;
	fast = (cd != O_NORMALMAP) && nrandom < 0;       /* :1139-1142 */

	if (fast) { /* :879-935 */
		ocol black = { 0, 0, 0 };
		int dmin = 0x7FFFFFFF, dmax = 0;
		c[0].r = 31; c[0].g = 63; c[0].b = 31;
		c[1].r = 0; c[1].g = 0; c[1].b = 0;
		if (dxt == O_DXT5) { ca[0] = rgba[3]; ca[1] = ca[0]; }
		for (x = 0; x < w; ++x)
			for (y = 0; y < h; ++y) {
				const uint8_t *p = &rgba[(x + y * iw) * 4];
				ocol t = { sc(p[0]), sc(p[1]), sc(p[2]) };
				uint8_t ta = p[3];
				int d;
				if (dxt == O_DXT1 && ta == 0) continue;
				d = cdist(cd, t, black);
				if (d > dmax) { dmax = d; c[1] = t; }
				if (d < dmin) { dmin = d; c[0] = t; }
				if (dxt == O_DXT5 && ta != 255) {
					if (ta > ca[1]) ca[1] = ta;
					if (ta < ca[0]) ca[0] = ta;
				}
			}
	} else { /* :936-1007 */
		int n = 0, m;
		for (x = 0; x < w; ++x)
			for (y = 0; y < h; ++y) {
				const uint8_t *p = &rgba[(x + y * iw) * 4];
				c[n].r = sc(p[0]); c[n].g = sc(p[1]); c[n].b = sc(p[2]);
				ca[n] = p[3];
				if (dxt == O_DXT1 && ca[n] == 0) continue;
				++n;
			}
		if (n == 0) {
			n = 1;
			c[0].r = 0; c[0].g = 0; c[0].b = 0;
			ca[0] = 0;
		}
		m = n;
		if (nrandom > 0) { /* :962-993 with the counter-based generator */
			ocol mins = c[0], maxs = c[0], len;
			int mina = (dxt == O_DXT5) ? ca[0] : 0, maxa = mina, lena;
			uint32_t k = 0;
			for (x = 1; x < n; ++x) {
				if (c[x].r < mins.r) mins.r = c[x].r;
				if (c[x].g < mins.g) mins.g = c[x].g;
				if (c[x].b < mins.b) mins.b = c[x].b;
				if (c[x].r > maxs.r) maxs.r = c[x].r;
				if (c[x].g > maxs.g) maxs.g = c[x].g;
				if (c[x].b > maxs.b) maxs.b = c[x].b;
				if (dxt == O_DXT5) {
					if (ca[x] < mina) mina = ca[x];
					if (ca[x] > maxa) maxa = ca[x];
				}
			}
			len.r = sc(maxs.r - mins.r + 1);
			len.g = sc(maxs.g - mins.g + 1);
			len.b = sc(maxs.b - mins.b + 1);
			lena = (dxt == O_DXT5) ? (maxa - mina + 1) : 0;
			for (x = 0; x < nrandom; ++x) {
				c[m].r = sc(mins.r + s2tc_prng_draw(bkey, k++) % len.r);
				c[m].g = sc(mins.g + s2tc_prng_draw(bkey, k++) % len.g);
				c[m].b = sc(mins.b + s2tc_prng_draw(bkey, k++) % len.b);
				if (dxt == O_DXT5)
					ca[m] = (uint8_t)(mina + s2tc_prng_draw(bkey, k++) % lena);
				++m;
			}
		} else if (n == 1) { /* "hack for last miplevel" :996-1001 */
			c[1] = c[0];
			m = n = 2;
		}
		reduce_colors(c, n, m, cd);
		if (dxt == O_DXT5) reduce_alpha_2fix(ca, n, m);
	}

	/* equal colors are BAD :1009-1027 */
	if (ceq(c[0], c[1])) {
		if (ceq(c[0], CMAX)) cdec(&c[1]); else cinc(&c[1]);
	}
	if (dxt == O_DXT5 && ca[0] == ca[1]) {
		if (ca[0] == 255) --ca[1]; else ++ca[1];
	}

	/* refine and pack :1029-1107 */
	if (dxt == O_DXT1) {
		cbits = color_refine(refine, cd, 1, rgba, iw, w, h, &c[0], &c[1]);
		pack565(out + 0, c[0]);
		pack565(out + 2, c[1]);
		out[4] = (uint8_t)cbits; out[5] = (uint8_t)(cbits >> 8);
		out[6] = (uint8_t)(cbits >> 16); out[7] = (uint8_t)(cbits >> 24);
	} else if (dxt == O_DXT3) {
		uint64_t abits = 0;
		int q;
		cbits = color_refine(refine, cd, 0, rgba, iw, w, h, &c[0], &c[1]);
		for (x = 0; x < w; ++x) for (y = 0; y < h; ++y) /* s2tc_dxt3_encode_alpha :862-870 */
			abits |= (uint64_t)rgba[(y * iw + x) * 4 + 3] << (4 * (y * 4 + x));
		for (q = 0; q < 8; ++q) out[q] = (uint8_t)(abits >> (8 * q));
		pack565(out + 8, c[0]);
		pack565(out + 10, c[1]);
		out[12] = (uint8_t)cbits; out[13] = (uint8_t)(cbits >> 8);
		out[14] = (uint8_t)(cbits >> 16); out[15] = (uint8_t)(cbits >> 24);
	} else {
		uint64_t abits;
		int q;
		cbits = color_refine(refine, cd, 0, rgba, iw, w, h, &c[0], &c[1]);
		abits = alpha_refine(refine, rgba, iw, w, h, &ca[0], &ca[1]);
		out[0] = ca[0];
		out[1] = ca[1];
		for (q = 0; q < 6; ++q) out[2 + q] = (uint8_t)(abits >> (8 * q));
		pack565(out + 8, c[0]);
		pack565(out + 10, c[1]);
		out[12] = (uint8_t)cbits; out[13] = (uint8_t)(cbits >> 8);
		out[14] = (uint8_t)(cbits >> 16); out[15] = (uint8_t)(cbits >> 24);
	}
	free(c);
	free(ca);
}

/* ---- rgb565_image and its dithers :1196-1465 ---- */
static inline int imax(int a, int b) { return b > a ? b : a; }
static inline int imin(int a, int b) { return b < a ? b : a; }

static int diffuse(int *diff, int src, int shift) /* :1198-1207 */
{
	const int maxval = (1 << (8 - shift)) - 1;
	int ret, loop;
	src += *diff;
	ret = imax(0, imin(src >> shift, maxval));
	loop = (ret << shift) | (ret >> (8 - 2 * shift));
	*diff = src - loop;
	return ret;
}
static int diffuse1(int *diff, int src) /* :1208-1216 */
{
	int ret;
	src += *diff;
	ret = (src >= 128);
	*diff = src - (ret ? 255 : 0);
	return ret;
}
static void fs_spread(int *thisrow, int *downrow, int err) /* tail of floyd/floyd1 :1227-1237 */
{
	int e7 = (err * 7 + 8) / 16, e3, e5;
	err -= e7;
	e3 = (err * 3 + 4) / 9;
	err -= e3;
	e5 = (err * 5 + 3) / 6;
	err -= e5;
	thisrow[2] += e7;
	downrow[0] += e3;
	downrow[1] += e5;
	downrow[2] += err;
}
static int floyd(int *thisrow, int *downrow, int src, int shift) /* :1218-1239 */
{
	const int maxval = (1 << (8 - shift)) - 1;
	int ret;
	src = (src << 4) | (src >> 4);
	src += thisrow[1];
	ret = imax(0, imin(src >> (shift + 4), maxval));
	fs_spread(thisrow, downrow, src - ret * 4095 / maxval);
	return ret;
}
static int floyd1(int *thisrow, int *downrow, int src) /* :1241-1261 */
{
	int ret;
	src = (src << 4) | (src >> 4);
	src += thisrow[1];
	ret = (src >= 2048);
	fs_spread(thisrow, downrow, src - (ret ? 4095 : 0));
	return ret;
}

/* srccomps != 3 is treated as 4 (:1455-1463); alphabits not 1/4 is 8 (:1437-1449);
 * dither outside NONE/FS is SIMPLE (:1419-1431). */
void oracle_rgb565_image(uint8_t *out, const uint8_t *rgba, int w, int h,
                         int srccomps, int alphabits, int dither)
{
	int x, y, sc4;
	if (srccomps != 3) srccomps = 4;
	if (alphabits != 1 && alphabits != 4) alphabits = 8;
	if (dither != O_DITHER_NONE && dither != O_DITHER_FLOYDSTEINBERG) dither = O_DITHER_SIMPLE;
	sc4 = srccomps;
#define SRC(x, y, c) rgba[((x) + (y) * w) * sc4 + (c)]
#define DST(x, y, c) out[((x) + (y) * w) * 4 + (c)]
	if (dither == O_DITHER_NONE) { /* :1269-1306 */
		for (y = 0; y < h; ++y) for (x = 0; x < w; ++x) {
			DST(x, y, 0) = SRC(x, y, 0) >> 3;
			DST(x, y, 1) = SRC(x, y, 1) >> 2;
			DST(x, y, 2) = SRC(x, y, 2) >> 3;
		}
		for (y = 0; y < h; ++y) for (x = 0; x < w; ++x) {
			if (srccomps == 4) {
				if (alphabits == 1) DST(x, y, 3) = SRC(x, y, 3) >> 7;
				else if (alphabits == 8) DST(x, y, 3) = SRC(x, y, 3);
				else DST(x, y, 3) = SRC(x, y, 3) >> (8 - alphabits);
			} else DST(x, y, 3) = (uint8_t)((1 << alphabits) - 1);
		}
	} else if (dither == O_DITHER_SIMPLE) { /* :1307-1349 */
		int dr = 0, dg = 0, db = 0, da = 0;
		for (y = 0; y < h; ++y) for (x = 0; x < w; ++x) {
			DST(x, y, 0) = (uint8_t)diffuse(&dr, SRC(x, y, 0), 3);
			DST(x, y, 1) = (uint8_t)diffuse(&dg, SRC(x, y, 1), 2);
			DST(x, y, 2) = (uint8_t)diffuse(&db, SRC(x, y, 2), 3);
		}
		for (y = 0; y < h; ++y) for (x = 0; x < w; ++x) {
			if (srccomps == 4) {
				if (alphabits == 1) DST(x, y, 3) = (uint8_t)diffuse1(&da, SRC(x, y, 3));
				else if (alphabits == 8) DST(x, y, 3) = SRC(x, y, 3);
				else DST(x, y, 3) = (uint8_t)diffuse(&da, SRC(x, y, 3), 8 - alphabits);
			} else DST(x, y, 3) = (uint8_t)((1 << alphabits) - 1);
		}
	} else { /* FLOYDSTEINBERG :1350-1412; one 6*(w+2) int buffer shared by the
	          * colour pass and the alpha pass, which starts on what the colour
	          * pass left in it (only the next-row half is cleared per row). */
		int pw = w + 2;
		int *buf = (int *)calloc((size_t)6 * pw, sizeof(int));
		for (y = 0; y < h; ++y) {
			int *tr = buf + ((y & 1) ? 3 : 0) * pw;
			int *dr = buf + ((y & 1) ? 0 : 3) * pw;
			memset(dr, 0, sizeof(int) * 3 * pw);
			for (x = 0; x < w; ++x) {
				DST(x, y, 0) = (uint8_t)floyd(&tr[x], &dr[x], SRC(x, y, 0), 3);
				DST(x, y, 1) = (uint8_t)floyd(&tr[pw + x], &dr[pw + x], SRC(x, y, 1), 2);
				DST(x, y, 2) = (uint8_t)floyd(&tr[2 * pw + x], &dr[2 * pw + x], SRC(x, y, 2), 3);
			}
		}
		if (srccomps == 4 && alphabits != 8) {
			for (y = 0; y < h; ++y) {
				int *tr = buf + (y & 1) * pw;
				int *dr = buf + !(y & 1) * pw;
				memset(dr, 0, sizeof(int) * pw);
				for (x = 0; x < w; ++x) {
					if (alphabits == 1) DST(x, y, 3) = (uint8_t)floyd1(&tr[x], &dr[x], SRC(x, y, 3));
					else DST(x, y, 3) = (uint8_t)floyd(&tr[x], &dr[x], SRC(x, y, 3), 8 - alphabits);
				}
			}
		} else {
			for (y = 0; y < h; ++y) for (x = 0; x < w; ++x)
				DST(x, y, 3) = (srccomps == 4) ? SRC(x, y, 3) : (uint8_t)((1 << alphabits) - 1);
		}
		free(buf);
	}
#undef SRC
#undef DST
}

/* ---- block loop of tx_compress_dxtn :246-294 ---- */
void oracle_encode_image_quantized(const uint8_t *q, int width, int height, int dxt,
                                   int cd, int nrandom, int refine, uint64_t seed,
                                   uint8_t *dest, int row0, int row1)
{
	int bw = (width + 3) / 4, bs = (dxt == O_DXT1) ? 8 : 16, bx, by;
	for (by = row0; by < row1; ++by)
		for (bx = 0; bx < bw; ++bx) {
			int i = bx * 4, j = by * 4;
			int nx = width > i + 3 ? 4 : width - i;
			int ny = height > j + 3 ? 4 : height - j;
			uint32_t bkey = s2tc_prng_block_key(seed, (uint32_t)(by * bw + bx));
			oracle_encode_block(dest + ((size_t)by * bw + bx) * bs, q + ((size_t)j * width + i) * 4,
			                    width, nx, ny, dxt, cd, nrandom, refine, bkey);
		}
}

/* tx_compress_dxtn :142-299 (settings passed in rather than read from env) */
int oracle_compress_image(int srccomps, int width, int height, const uint8_t *src,
                          unsigned fmt, uint8_t *dest, int dstRowStride,
                          int cd, int nrandom, int refine, int dither, uint64_t seed)
{
	int dxt, alphabits, k, bs, bw, bh, by, rowdiff;
	uint8_t *q, *tmp, *blk;
	switch (fmt) {
	case 0x83F0: case 0x83F1: dxt = O_DXT1; alphabits = 1; k = 2; bs = 8; break;
	case 0x83F2: dxt = O_DXT3; alphabits = 4; k = 4; bs = 16; break;
	case 0x83F3: dxt = O_DXT5; alphabits = 8; k = 4; bs = 16; break;
	default: return -1;
	}
	q = (uint8_t *)malloc((size_t)width * height * 4 + 1);
	oracle_rgb565_image(q, src, width, height, srccomps, alphabits, dither);
	bw = (width + 3) / 4;
	bh = (height + 3) / 4;
	tmp = (uint8_t *)malloc((size_t)bw * bh * bs + 1);
	oracle_encode_image_quantized(q, width, height, dxt, cd, nrandom, refine, seed, tmp, 0, bh);
	/* dstRowDiff :243, :261, :279 */
	rowdiff = dstRowStride >= width * k ? dstRowStride - ((width + 3) & ~3) * k : 0;
	blk = dest;
	for (by = 0; by < bh; ++by) {
		memcpy(blk, tmp + (size_t)by * bw * bs, (size_t)bw * bs);
		blk += bw * bs + rowdiff;
	}
	free(tmp);
	free(q);
	return 0;
}

/* ---- fetch_2d_texel_* :35-140 ---- */
static void expand565(unsigned c, uint8_t *t)
{
	t[0] = (uint8_t)((c >> 11) & 0x1F); t[0] = (uint8_t)((t[0] << 3) | (t[0] >> 2));
	t[1] = (uint8_t)((c >> 5) & 0x3F);  t[1] = (uint8_t)((t[1] << 2) | (t[1] >> 4));
	t[2] = (uint8_t)(c & 0x1F);         t[2] = (uint8_t)((t[2] << 3) | (t[2] >> 2));
}
void oracle_fetch_texel(unsigned fmt, int srcRowStride, const uint8_t *pix, int i, int j, uint8_t *t)
{
	int bpb = (fmt == 0x83F0 || fmt == 0x83F1) ? 8 : 16;
	const uint8_t *blk = pix + ((size_t)((srcRowStride + 3) >> 2) * (j >> 2) + (i >> 2)) * bpb;
	const uint8_t *cb = (bpb == 8) ? blk : blk + 8;
	unsigned c = cb[0] + 256u * cb[1], c1 = cb[2] + 256u * cb[3];
	int b = (cb[4 + (j & 3)] >> (2 * (i & 3))) & 3;
	int alpha = 255;
	if (bpb == 8) {
		if (b == 1) c = c1;
		else if (b == 3 && c1 >= c) { c = 0; if (fmt == 0x83F1) alpha = 0; }
		else if (b >= 2) { if ((i ^ j) & 1) c = c1; }
	} else {
		if (b == 1) c = c1;
		else if (b >= 2) { if ((i ^ j) & 1) c = c1; }
	}
	expand565(c, t);
	if (fmt == 0x83F0 || fmt == 0x83F1) { t[3] = (uint8_t)alpha; return; }
	if (fmt == 0x83F2) {
		int a = (blk[(j & 3) * 2 + ((i & 3) >> 1)] >> (4 * (i & 1))) & 0x0F;
		t[3] = (uint8_t)(a | (a << 4));
	} else {
		unsigned a = blk[0], a1 = blk[1];
		int abit = ((j & 3) * 4 + (i & 3)) * 3, ab = 0, q;
		for (q = 0; q < 3; ++q)
			if (blk[2 + ((abit + q) >> 3)] & (1 << ((abit + q) & 7))) ab |= 1 << q;
		/* switch with fall-through from 6 into 7 (:130-137) */
		switch (ab) {
		case 0: break;
		case 1: a = a1; break;
		case 6: if (a1 >= a) { a = 0; break; } /* fall through */
		case 7: if (a1 >= a) { a = 255; break; } /* fall through */
		default: if ((i ^ j) & 1) a = a1; break;
		}
		t[3] = (uint8_t)a;
	}
}

/* every texel of a width x height image through oracle_fetch_texel, row-major RGBA8 */
void oracle_decode_image(unsigned fmt, int width, int height, const uint8_t *pix, uint8_t *out)
{
	int i, j;
	for (j = 0; j < height; ++j)
		for (i = 0; i < width; ++i) oracle_fetch_texel(fmt, width, pix, i, j, out + ((size_t)j * width + i) * 4);
}

/* ---- total colour distance of an encoding (test metric) ---- */
int64_t oracle_colordist_error(const uint8_t *q, int width, int height, int dxt, int cd, const uint8_t *blocks)
{
	const int bw = (width + 3) / 4, bs = dxt == O_DXT1 ? 8 : 16;
	int64_t total = 0;
	int i, j;
	for (j = 0; j < height; ++j)
		for (i = 0; i < width; ++i) {
			const uint8_t *t = q + ((size_t)j * width + i) * 4;
			const uint8_t *cb = blocks + ((size_t)(j >> 2) * bw + (i >> 2)) * bs + (dxt == O_DXT1 ? 0 : 8);
			const unsigned c0 = cb[0] + 256u * cb[1], c1 = cb[2] + 256u * cb[3];
			const int b = (cb[4 + (j & 3)] >> (2 * (i & 3))) & 3;
			unsigned c = c0;
			ocol tc, dc;
			if (dxt == O_DXT1 && t[3] == 0) continue;
			if (b == 1) c = c1;
			else if (dxt == O_DXT1 && b == 3 && c1 >= c0) c = 0;
			else if (b >= 2 && ((i ^ j) & 1)) c = c1;
			tc.r = sc(t[0]); tc.g = sc(t[1]); tc.b = sc(t[2]);
			dc.r = sc((int)(c >> 11) & 31); dc.g = sc((int)(c >> 5) & 63); dc.b = sc((int)c & 31);
			total += cdist(cd, tc, dc);
		}
	return total;
}

/* ---- Image_MipReduce32 toward 1x1 (s2tc_compress.c:427-493) ---- */
void oracle_mip_reduce(const uint8_t *in, uint8_t *out, int *width, int *height)
{
	const uint8_t *inrow = in;
	int x, y, nextrow = *width * 4, c;
	if (*width > 1) {
		*width >>= 1;
		if (*height > 1) {
			*height >>= 1;
			for (y = 0; y < *height; y++, inrow += nextrow * 2)
				for (in = inrow, x = 0; x < *width; x++, out += 4, in += 8)
					for (c = 0; c < 4; ++c)
						out[c] = (uint8_t)((in[c] + in[4 + c] + in[nextrow + c] + in[nextrow + 4 + c]) >> 2);
		} else {
			for (y = 0; y < *height; y++, inrow += nextrow)
				for (in = inrow, x = 0; x < *width; x++, out += 4, in += 8)
					for (c = 0; c < 4; ++c)
						out[c] = (uint8_t)((in[c] + in[4 + c]) >> 1);
		}
	} else if (*height > 1) {
		*height >>= 1;
		for (y = 0; y < *height; y++, inrow += nextrow * 2)
			for (in = inrow, x = 0; x < *width; x++, out += 4, in += 4)
				for (c = 0; c < 4; ++c)
					out[c] = (uint8_t)((in[c] + in[nextrow + c]) >> 1);
	}
}

/* ---- s2tc_from_s3tc: s2tc_from_s3tc.cpp:77-190, 254-263 ----
 *
 * Restated per texel index rather than with the reference's word-wide bit
 * operations.  Texel k = 4y + x of the block; "checker" is (x + y) & 1, the
 * pattern in which the reference sends an index that S2TC cannot express to
 * one of the two endpoints.
 *
 * Colour block (2-bit indices, words c0 = bytes 0-1, c1 = bytes 2-3):
 *   keep3 (c1 >= c0 in a DXT1 block: three colours + transparent, kept):
 *     0 -> 0, 1 -> 1, 2 -> checker, 3 -> 3;
 *   no3   (every other case): 0 -> 0, 1 -> 1, 2 -> checker, 3 -> checker;
 *   then, to give S2TC's endpoint order, the endpoints swap and indices
 *   0 <-> 1 flip: in a DXT1 block when c1 < c0, in a DXT3/5 colour block
 *   when c1 >= c0.
 * DXT5 alpha block (3-bit indices, a0 = byte 0, a1 = byte 1):
 *   a1 >= a0 (six-value mode): 0, 1, 6, 7 kept, 2..5 -> checker;
 *   a1 <  a0 (eight-value mode): 0, 1 kept, 2..7 -> checker, then the
 *   endpoints swap and indices 0 <-> 1 flip. */
static int fs_checker(int k) { return ((k & 3) + (k >> 2)) & 1; }

static void fs_colour(uint8_t *b, int dxt1)
{
	const unsigned c0 = b[0] | (b[1] << 8), c1 = b[2] | (b[3] << 8);
	const uint32_t idx = b[4] | (b[5] << 8) | (b[6] << 16) | ((uint32_t)b[7] << 24);
	const int keep3 = dxt1 && c1 >= c0;
	const int swap = dxt1 ? c1 < c0 : c1 >= c0;
	uint32_t out = 0;
	for (int k = 0; k < 16; ++k) {
		int v = (idx >> (2 * k)) & 3;
		if (v == 2 || (v == 3 && !keep3)) v = fs_checker(k);
		if (swap) v ^= 1;   /* v is 0 or 1 here: a swap never meets a kept 3 */
		out |= (uint32_t)v << (2 * k);
	}
	if (swap) {
		b[0] = (uint8_t)c1; b[1] = (uint8_t)(c1 >> 8);
		b[2] = (uint8_t)c0; b[3] = (uint8_t)(c0 >> 8);
	}
	for (int i = 0; i < 4; ++i) b[4 + i] = (uint8_t)(out >> (8 * i));
}

static void fs_alpha(uint8_t *b)
{
	const int a0 = b[0], a1 = b[1];
	uint64_t idx = 0, out = 0;
	for (int i = 0; i < 6; ++i) idx |= (uint64_t)b[2 + i] << (8 * i);
	const int six = a1 >= a0;
	for (int k = 0; k < 16; ++k) {
		int v = (int)((idx >> (3 * k)) & 7);
		if (six ? (v >= 2 && v <= 5) : v >= 2) v = fs_checker(k);
		if (!six) v ^= 1;
		out |= (uint64_t)v << (3 * k);
	}
	if (!six) {
		b[0] = (uint8_t)a1;
		b[1] = (uint8_t)a0;
	}
	for (int i = 0; i < 6; ++i) b[2 + i] = (uint8_t)(out >> (8 * i));
}

void oracle_from_s3tc_block(uint8_t *block, int dxt)
{
	if (dxt == O_DXT1) {
		fs_colour(block, 1);
		return;
	}
	fs_colour(block + 8, 0);
	if (dxt == O_DXT5) fs_alpha(block);
}
