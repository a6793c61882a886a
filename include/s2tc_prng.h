/*
 * s2tc_prng.h - counter-based random candidate generator for S2TC.
 *
 * The reference draws its extra candidate colors with the process-global
 * libc rand() stream (s2tc_algorithm.cpp:984-992): per block and per extra
 * candidate x = 0..nrandom-1 it draws, in this order,
 *     r = mins.r + rand() % len.r
 *     g = mins.g + rand() % len.g
 *     b = mins.b + rand() % len.b
 *     a = mina   + rand() % lena        (DXT5 only)
 * i.e. exactly nrandom * (3 or 4) draws per block, always.  A sequential
 * global stream cannot be reproduced by a parallel encoder, so this library
 * replaces rand() by a pure function of (seed, block index, draw index):
 *
 *     draw(seed, block, k) = mix32(block_key(seed, block) + k * 0x9E3779B9) & 0x7FFFFFFF
 *
 * which, like rand(), is a non-negative 31-bit int.  Draw k of a block is
 * candidate k / dpc, channel k % dpc, with dpc = 3 (DXT1/DXT3) or 4 (DXT5).
 * The block index is the raster-order index of the 4x4 block inside one
 * image (or mip level): (y / 4) * ceil(width / 4) + x / 4.
 *
 * The same definitions are used by the GPU kernels, by the CPU oracle and
 * by the rand() interposer linked into the compiled reference, so all three
 * produce bit-identical blocks for nrandom > 0.
 *
 * mix32 is the "lowbias32" integer finaliser (C. Wellons, public domain
 * hash-prospector result): xorshift-multiply rounds with the constants
 * 0x7feb352d and 0x846ca68b.
 */
#ifndef S2TC_PRNG_H
#define S2TC_PRNG_H

#include <stdint.h>

#if defined(__HIP__) || defined(__HIPCC__)
#define S2TC_PRNG_FN __host__ __device__ static inline
#else
#define S2TC_PRNG_FN static inline
#endif

S2TC_PRNG_FN uint32_t s2tc_mix32(uint32_t x)
{
	x ^= x >> 16;
	x *= 0x7feb352dU;
	x ^= x >> 15;
	x *= 0x846ca68bU;
	x ^= x >> 16;
	return x;
}

/* Key of one 4x4 block: a function of the 64-bit seed and the block index. */
S2TC_PRNG_FN uint32_t s2tc_prng_block_key(uint64_t seed, uint32_t block)
{
	uint32_t k = s2tc_mix32((uint32_t)seed ^ 0x9E3779B9U);
	k = s2tc_mix32(k ^ (uint32_t)(seed >> 32));
	return s2tc_mix32(k ^ block);
}

/* Draw k (0-based, in the reference's draw order) of the block with key bkey. */
S2TC_PRNG_FN int s2tc_prng_draw(uint32_t bkey, uint32_t k)
{
	return (int)(s2tc_mix32(bkey + k * 0x9E3779B9U) & 0x7FFFFFFFU);
}

/* Seed of one mip level of one texture in a batch: used by the batch and
 * serving entry points so that every (texture, level) has its own stream. */
S2TC_PRNG_FN uint64_t s2tc_prng_level_seed(uint64_t texture_seed, uint32_t level)
{
	uint32_t lo = s2tc_mix32((uint32_t)texture_seed ^ (level * 0x85EBCA6BU));
	uint32_t hi = s2tc_mix32((uint32_t)(texture_seed >> 32) ^ lo ^ level);
	return ((uint64_t)hi << 32) | lo;
}

#endif
