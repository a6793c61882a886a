// encode_all16.h - MODE_NORMAL with the block's own colours only
// (nrandom <= 0: "all 16 candidates", or NORMALMAP without random colours),
// one lane per block.
//
// With at most 16 candidates the pair scan of reduce_colors_inplace
// (s2tc_algorithm.cpp:363-414) is 120 pairs x 16 texels, small enough for
// one lane: the 120 distances of the upper triangle stay in registers and
// the lane folds the pairs in the reference's (i, j) order with the
// reference's own update rule "bestsum < 0 || sum < bestsum", so the
// int32-wrapped negative sums of SRGB need no special handling.
//
// Candidates are not compacted: slot k is texel (x = k >> 2, y = k & 3) and
// takes part iff it is inside the block and (for DXT1) not transparent.
// Dropping unused slots keeps the order of the rest, so pair order, sums and
// tie-breaks equal those over the reference's compacted list.
#pragma once

#include "encode_fast.h"

namespace s2tc {

// Upper-triangle index of (i, j), i < j < 16.
__host__ __device__ constexpr int tri(int i, int j) { return i * 16 - i * (i + 1) / 2 + (j - i - 1); }

// Argmin of the pair scan; D(i, k) = dists[i][k] of the reference (symmetric
// storage of the square: dists[i][k] = dist(c[min], c[max]), diagonal 0).
// use: bit k set if slot k is a candidate.  Returns the pair as (bi, bj).
template <bool NONNEG>
S2TC_D void fold_pairs(const int (&D)[120], uint32_t use, int &bi, int &bj)
{
	int bestsum = -1;
	bi = 0;
	bj = 1;
#pragma unroll
	for (int i = 0; i < 16; ++i) {
#pragma unroll
		for (int j = i + 1; j < 16; ++j) {
			int s = 0;
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				const int di = k == i ? 0 : D[k < i ? tri(k, i) : tri(i, k)];
				const int dj = k == j ? 0 : D[k < j ? tri(k, j) : tri(j, k)];
				s = (int)((uint32_t)s + (uint32_t)(di < dj ? di : dj));
			}
			const bool ok = ((use >> i) & 1u) && ((use >> j) & 1u);
			const bool take = ok && (NONNEG ? (uint32_t)s < (uint32_t)bestsum : (bestsum < 0 || s < bestsum));
			bestsum = take ? s : bestsum;
			bi = take ? i : bi;
			bj = take ? j : bj;
		}
	}
}

// As fold_pairs for reduce_colors_inplace_2fixpoints (:415-478): every term
// is also bounded by f[k] = min(dist(0, a_k), dist(255, a_k)).  Alpha
// distances are never negative.
S2TC_D void fold_pairs_2fix(const int (&A)[120], const int (&f)[16], uint32_t use, int &bi, int &bj)
{
	uint32_t bestsum = 0xffffffffu;
	bi = 0;
	bj = 1;
#pragma unroll
	for (int i = 0; i < 16; ++i) {
#pragma unroll
		for (int j = i + 1; j < 16; ++j) {
			uint32_t s = 0;
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				const int di = k == i ? 0 : A[k < i ? tri(k, i) : tri(i, k)];
				const int dj = k == j ? 0 : A[k < j ? tri(k, j) : tri(j, k)];
				int t = di < dj ? di : dj;
				t = t < f[k] ? t : f[k];
				s += (uint32_t)t;
			}
			const bool take = ((use >> i) & 1u) && ((use >> j) & 1u) && s < bestsum;
			bestsum = take ? s : bestsum;
			bi = take ? i : bi;
			bj = take ? j : bj;
		}
	}
}

template <int CD, int DXT>
S2TC_D BlockOut encode_block_all16(const uint32_t (&tx)[16], int w, int h, int refine)
{
	typedef Dist<CD> Dd;
	typedef typename Dd::F F;
	uint32_t valid = 0, use = 0;
#pragma unroll
	for (int k = 0; k < 16; ++k) {
		const int x = k >> 2, y = k & 3;
		const bool v = x < w && y < h;
		valid |= (v ? 1u : 0u) << k;
		use |= ((v && !(DXT == DXT1 && texel_a(tx[k]) == 0)) ? 1u : 0u) << k;
	}
	const int n = __builtin_popcount(use);

	uint32_t c0, c1;
	int a0 = 0, a1 = 0;
	if (n <= 1) {
		// one candidate (or none: black) duplicated by the "hack for last
		// miplevel" (:952-959, :996-1001); its pair has sum 0
		const int k0 = use ? __builtin_ctz(use) : 0;
		uint32_t c = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) c = k == k0 ? texel565(tx[k]) : c;
		c0 = c1 = use ? c : 0u;
		if (DXT == DXT5) {  // ca = {alpha of the texel, 0}: the uninitialised ca[1] pinned to 0
			int a = 0;
#pragma unroll
			for (int k = 0; k < 16; ++k) a = k == k0 ? texel_a(tx[k]) : a;
			a0 = a;
			a1 = 0;
		}
	} else {
		int bi, bj;
		{
			// the distances of the upper triangle; texels outside the
			// candidate set contribute nothing to any sum
			int D[120];
			{
				F f[16];
#pragma unroll
				for (int k = 0; k < 16; ++k) f[k] = Dd::featq(tx[k]);
#pragma unroll
				for (int i = 0; i < 16; ++i)
#pragma unroll
					for (int j = i + 1; j < 16; ++j)
						D[tri(i, j)] = (((use >> i) & (use >> j)) & 1u) ? Dd::d(f[i], f[j]) : 0;
			}
			fold_pairs<IsSymmetric<CD>::value>(D, use, bi, bj);
		}
		uint32_t v0 = 0, v1 = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			v0 = k == bi ? texel565(tx[k]) : v0;
			v1 = k == bj ? texel565(tx[k]) : v1;
		}
		c0 = v0;
		c1 = v1;
		if (DXT == DXT5) {
			int ai, aj;
			{
				int A[120], fx[16];
#pragma unroll
				for (int k = 0; k < 16; ++k) {
					const int ak = texel_a(tx[k]);
					const int f0 = ak * ak, f1 = (255 - ak) * (255 - ak);
					fx[k] = ((valid >> k) & 1u) ? (f1 < f0 ? f1 : f0) : 0;
				}
#pragma unroll
				for (int i = 0; i < 16; ++i)
#pragma unroll
					for (int j = i + 1; j < 16; ++j)
						A[tri(i, j)] =
						    (((valid >> i) & (valid >> j)) & 1u) ? adist(texel_a(tx[i]), texel_a(tx[j])) : 0;
				fold_pairs_2fix(A, fx, valid, ai, aj);
			}
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				a0 = k == ai ? texel_a(tx[k]) : a0;
				a1 = k == aj ? texel_a(tx[k]) : a1;
			}
		}
	}
	TexFeat<CD> f;
	f.init(tx);
	const Masks mk = block_masks<DXT>(tx, w, h);
	c1 = fix_equal565(c0, c1);
	BlockOut o;
	if (DXT == DXT1) {
		const uint32_t bits = color_refine_lane<CD, true, false>(tx, f, mk, refine, c0, c1);
		o.w[0] = c0 | (c1 << 16);
		o.w[1] = bits;
		o.w[2] = o.w[3] = 0;
	} else if (DXT == DXT3) {
		const uint32_t bits = color_refine_lane<CD, false, false>(tx, f, mk, refine, c0, c1);
		uint64_t ab = 0;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const int x = k >> 2, y = k & 3;
			ab |= ((valid >> k) & 1u) ? (uint64_t)texel_a(tx[k]) << (4 * (y * 4 + x)) : 0ull;
		}
		o.w[0] = (uint32_t)ab;
		o.w[1] = (uint32_t)(ab >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	} else {
		a1 = fix_equal_a(a0, a1);
		const uint32_t bits = color_refine_lane<CD, false, false>(tx, f, mk, refine, c0, c1);
		const uint64_t ab = alpha_refine_lane<false>(tx, mk, refine, a0, a1);
		const uint64_t lo = (uint64_t)(a0 & 255) | ((uint64_t)(a1 & 255) << 8) | (ab << 16);
		o.w[0] = (uint32_t)lo;
		o.w[1] = (uint32_t)(lo >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	}
	return o;
}

}  // namespace s2tc
