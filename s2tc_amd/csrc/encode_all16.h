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


// D(i, k) = dists[i][k] of the reference, stored as the upper triangle of the
// symmetric square (diagonal 0).
S2TC_D int dpair(const int (&D)[120], int i, int k) { return k == i ? 0 : D[k < i ? tri(k, i) : tri(i, k)]; }

// Argmin of the pair scan of reduce_colors_inplace (:363-414) for bounded
// non-negative distances, as one unsigned min over keys
// (sum << 8 | i << 4 | j): the smallest sum, then the first pair in the
// reference's (i, j) order, which is its "sum < bestsum" rule.  Terms with
// k == i or k == j are min(0, .) = 0 and are skipped.  A pair with a slot
// outside the candidate set starts at 2^23, above every real sum (< 2^23),
// so it never wins.  F: the 2-fixpoint variant (:415-478), every term also
// bounded by f[k].  Returns (i << 4 | j).
template <bool F>
S2TC_D uint32_t fold_pairs_key(const int (&D)[120], const int (&f)[16], uint32_t use)
{
	uint32_t pen[16];
#pragma unroll
	for (int k = 0; k < 16; ++k) pen[k] = ((use >> k) & 1u) ? 0u : (1u << 23);
	uint32_t best = 0xffffffffu;
#pragma unroll
	for (int i = 0; i < 16; ++i) {
#pragma unroll
		for (int j = i + 1; j < 16; ++j) {
			uint32_t s = pen[i] | pen[j];
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				if (k == i || k == j) continue;
				int t = min(dpair(D, i, k), dpair(D, j, k));
				if (F) t = min(t, f[k]);
				s += (uint32_t)t;
			}
			best = min(best, (s << 8) | (uint32_t)(i << 4 | j));
		}
	}
	return best & 255u;
}

// Argmin of the pair scan for any distance, with the reference's own update
// rule "bestsum < 0 || sum < bestsum" (int32-wrapped negative SRGB sums
// included).  use: bit k set if slot k is a candidate.  Returns (i << 4 | j).
template <bool NONNEG>
S2TC_D uint32_t fold_pairs(const int (&D)[120], uint32_t use)
{
	int bestsum = -1;
	uint32_t best = 1u;
#pragma unroll
	for (int i = 0; i < 16; ++i) {
#pragma unroll
		for (int j = i + 1; j < 16; ++j) {
			int s = 0;
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				if (NONNEG && (k == i || k == j)) continue;
				const int di = dpair(D, i, k), dj = dpair(D, j, k);
				s = (int)((uint32_t)s + (uint32_t)(di < dj ? di : dj));
			}
			const bool ok = ((use >> i) & 1u) && ((use >> j) & 1u);
			const bool take = ok && (NONNEG ? (uint32_t)s < (uint32_t)bestsum : (bestsum < 0 || s < bestsum));
			bestsum = take ? s : bestsum;
			best = take ? (uint32_t)(i << 4 | j) : best;
		}
	}
	return best;
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
	// a full DXT5 block with one alpha needs no alpha search
	const bool auni = DXT == DXT5 && valid == 0xffffu && alpha_uniform(tx);

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
		// Every candidate of one colour: whichever pair the scan picks, both of
		// its colours are that one (:996-1001 takes c[0], c[1] of the pair), so
		// the scan is skipped; a wave of such blocks skips it altogether.
		uint32_t lo = 0xffffffffu, hi = 0u;
#pragma unroll
		for (int k = 0; k < 16; ++k) {
			const uint32_t cw = tx[k] & 0xffffffu;
			const bool u = (use >> k) & 1u;
			lo = u && cw < lo ? cw : lo;
			hi = u && cw > hi ? cw : hi;
		}
		if (lo == hi) {
			c0 = c1 = texel565(lo);
		} else {
			int bi, bj;
			{
				// the distances of the upper triangle; texels outside the
				// candidate set contribute nothing to any sum
				int D[120];
				{
					// every distance is computed (no per-pair branch) and then
					// cleared where a slot is not a candidate
					F f[16];
					uint32_t um[16];
#pragma unroll
					for (int k = 0; k < 16; ++k) {
						f[k] = Dd::featq(tx[k]);
						um[k] = 0u - ((use >> k) & 1u);
					}
#pragma unroll
					for (int i = 0; i < 16; ++i)
#pragma unroll
						for (int j = i + 1; j < 16; ++j)
							D[tri(i, j)] = (int)((uint32_t)Dd::d(f[i], f[j]) & um[i] & um[j]);
				}
				__builtin_amdgcn_sched_barrier(0);
				uint32_t p;
				if constexpr (DistBound<CD>::value != 0) {
					const int none[16] = {};
					p = fold_pairs_key<false>(D, none, use);
				} else {
					p = fold_pairs<IsSymmetric<CD>::value>(D, use);
				}
				bi = (int)(p >> 4);
				bj = (int)(p & 15u);
			}
			uint32_t v0 = 0, v1 = 0;
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				v0 = k == bi ? texel565(tx[k]) : v0;
				v1 = k == bj ? texel565(tx[k]) : v1;
			}
			c0 = v0;
			c1 = v1;
		}
		if (DXT == DXT5 && !auni) {
			// keep the alpha triangle out of the colour triangle's live range
			__builtin_amdgcn_sched_barrier(0);
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
				const uint32_t p = fold_pairs_key<true>(A, fx, valid);
				ai = (int)(p >> 4);
				aj = (int)(p & 15u);
			}
#pragma unroll
			for (int k = 0; k < 16; ++k) {
				a0 = k == ai ? texel_a(tx[k]) : a0;
				a1 = k == aj ? texel_a(tx[k]) : a1;
			}
		}
	}
	__builtin_amdgcn_sched_barrier(0);
	// NORMALMAP DXT5: reload the texel features from the table here rather
	// than keep the triangle's 48 VGPRs of them live across the alpha scan
	if constexpr (!Dist<CD>::kCheap && DXT == DXT5) asm volatile("" ::: "memory");
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
		const uint32_t bits = color_refine_lane<CD, false, false>(tx, f, mk, refine, c0, c1);
		uint64_t lo;
		if (auni) {
			lo = alpha_uniform_half(texel_a(tx[0]));
		} else {
			a1 = fix_equal_a(a0, a1);
			const uint64_t ab = alpha_refine_fast<false>(tx, mk.valid, refine, a0, a1);
			lo = (uint64_t)(a0 & 255) | ((uint64_t)(a1 & 255) << 8) | (ab << 16);
		}
		o.w[0] = (uint32_t)lo;
		o.w[1] = (uint32_t)(lo >> 32);
		o.w[2] = c0 | (c1 << 16);
		o.w[3] = bits;
	}
	return o;
}

}  // namespace s2tc
