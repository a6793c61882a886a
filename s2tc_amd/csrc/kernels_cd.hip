// kernels_cd.hip - encoder kernels for one colour distance mode.
//
// Compiled once per mode with -DS2TC_CD=<ColorDistMode value>; each object
// exports s2tc_kernels_cd<N>() returning its kernel table.  Splitting by mode
// keeps each translation unit small and lets the build compile them in
// parallel.
#include "encode_all16.h"
#include "encode_normal.h"
#include "encode_lin.h"
#include "encode_quad.h"
#include "kernel_table.h"

#ifndef S2TC_CD
#error "S2TC_CD must be defined"
#endif

namespace s2tc {

// quantize one RGBA8 texel with DITHER_NONE (s2tc_algorithm.cpp:1269-1306)
template <int DXT> S2TC_D uint32_t quant_none(uint32_t p)
{
	uint32_t r = (p & 0xffu) >> 3, g = ((p >> 8) & 0xffu) >> 2, b = ((p >> 16) & 0xffu) >> 3, a = p >> 24;
	a = DXT == DXT1 ? a >> 7 : DXT == DXT3 ? a >> 4 : a;
	return r | (g << 8) | (b << 16) | (a << 24);
}
S2TC_D uint32_t quant_none_rt(uint32_t p, int dxt)
{
	return dxt == DXT1 ? quant_none<DXT1>(p) : dxt == DXT3 ? quant_none<DXT3>(p) : quant_none<DXT5>(p);
}

// smallest of the 16 source words: its top byte is the smallest alpha
S2TC_D uint32_t tx_min16(const uint32_t (&tx)[16])
{
	uint32_t m = tx[0];
#pragma unroll
	for (int k = 1; k < 16; ++k) m = min(m, tx[k]);
	return m;
}

// Load the 4x4 block (bx, by) into tx[x*4+y]; texels outside w x h are 0.
// Full blocks of 16-byte aligned rows (vec: width % 4 == 0) take four
// unconditional 16-byte loads.
template <int DXT, bool RAW>
S2TC_D void load_block(const uint32_t *__restrict__ src, int width, int bx, int by, int w, int h, bool vec,
                       uint32_t (&tx)[16])
{
	const uint32_t *base = src + (size_t)(4 * by) * (size_t)width + 4 * bx;
	if (vec && w == 4 && h == 4) {
#pragma unroll
		for (int y = 0; y < 4; ++y) {
			const uint4 q = *(const uint4 *)(base + y * width);
			tx[y] = q.x;
			tx[4 + y] = q.y;
			tx[8 + y] = q.z;
			tx[12 + y] = q.w;
		}
	} else {
#pragma unroll
		for (int k = 0; k < 16; ++k) tx[k] = 0;
#pragma unroll
		for (int y = 0; y < 4; ++y) {
			if (y < h) {
				const uint32_t *row = base + y * width;
#pragma unroll
				for (int x = 0; x < 4; ++x)
					if (x < w) tx[x * 4 + y] = row[x];
			}
		}
	}
	if (RAW) {
#pragma unroll
		for (int k = 0; k < 16; ++k) tx[k] = quant_none<DXT>(tx[k]);
	}
}

template <int DXT> S2TC_D void store_block(uint8_t *dst, size_t b, const BlockOut &o)
{
	if (DXT == DXT1)
		*(uint2 *)(dst + b * 8) = make_uint2(o.w[0], o.w[1]);
	else
		*(uint4 *)(dst + b * 16) = make_uint4(o.w[0], o.w[1], o.w[2], o.w[3]);
}

// Single-image kernels run 64 x 4 lanes per workgroup on a 2-D grid: lane
// (x, y) takes block (64 blockIdx.x + x, by0 + 4 blockIdx.y + y), so each row
// of 64 lanes reads 64 consecutive blocks (1 KiB per texel row) and no lane
// divides by the block-row width.
#define S2TC_IMG_BLOCK_COORDS                                                        \
	const int bw = (width + 3) >> 2, bh = (height + 3) >> 2;                         \
	const int bx = (int)(blockIdx.x * 64 + threadIdx.x);                             \
	const int by = by0 + (int)(blockIdx.y * 4 + threadIdx.y);                        \
	if (bx >= bw || by >= bh) return;                                                \
	const size_t b = (size_t)by * (size_t)bw + (size_t)bx;                          \
	const int w = min(4, width - 4 * bx), h = min(4, height - 4 * by);

// Quick mode over one image: one lane per block.
template <int CD, int DXT, bool RAW>
__global__ __launch_bounds__(256, 4) void k_fast(const uint32_t *__restrict__ src, int width, int height,
                                                 uint8_t *__restrict__ dst, int refine, int vec, int by0)
{
	S2TC_IMG_BLOCK_COORDS
	uint32_t tx[16];
	BlockOut o;
	if constexpr (QuadW<CD>::value) {
		load_block<DXT, false>(src, width, bx, by, w, h, vec != 0, tx);
		const uint32_t *blk = src + (size_t)(4 * by) * (size_t)width + 4 * bx;
		o = encode_block_quad_any<CD, DXT, RAW>(tx, blk, width, w, h, refine);
	} else {
		bool lin = false;
		if constexpr (LinW<CD>::value) lin = refine != REFINE_LOOP;
		if (lin) {
			load_block<DXT, false>(src, width, bx, by, w, h, vec != 0, tx);
			const uint32_t *blk = src + (size_t)(4 * by) * (size_t)width + 4 * bx;
			if constexpr (LinW<CD>::value) o = encode_block_lin_any<CD, DXT, RAW>(tx, blk, width, w, h, refine);
		} else {
			load_block<DXT, RAW>(src, width, bx, by, w, h, vec != 0, tx);
			o = encode_block_fast<CD, DXT>(tx, w, h, refine);
		}
	}
	store_block<DXT>(dst, b, o);
}

// Quick mode over the full 4x4 blocks of an image whose rows are 16-byte
// aligned (width % 4 == 0): blocks (bx, by) with bx < nbx, by0 <= by < nby.
// Without partial-block code the kernel needs far fewer registers, so more
// waves hide the loads; the edge blocks go to k_fast.
template <int CD, int DXT, bool RAW, int REFINE>
__global__ __launch_bounds__(256) void k_fast_interior(const uint32_t *__restrict__ src, int width, int nbx, int nby,
                                                       uint8_t *__restrict__ dst, int by0)
{
	const int refine = REFINE;
	const int bx = (int)(blockIdx.x * 64 + threadIdx.x);
	const int by = by0 + (int)(blockIdx.y * 4 + threadIdx.y);
	if (bx >= nbx || by >= nby) return;
	const size_t b = (size_t)by * (size_t)((width + 3) >> 2) + (size_t)bx;
	const uint32_t *blk = src + (size_t)(4 * by) * (size_t)width + 4 * bx;
	uint32_t tx[16];
#pragma unroll
	for (int y = 0; y < 4; ++y) {
		const uint4 q = *(const uint4 *)(blk + (size_t)y * (size_t)width);
		tx[y] = q.x;
		tx[4 + y] = q.y;
		tx[8 + y] = q.z;
		tx[12 + y] = q.w;
	}
	BlockOut o;
	if constexpr (QuadW<CD>::value) {
		if (DXT == DXT1 && __all((int)(RAW ? tx_min16(tx) >> 31 : tx_min16(tx) >> 24) != 0)) {
			opaque_to_compiler(tx);
			o = encode_block_quad<CD, DXT, true, RAW, true>(tx, blk, width, 4, 4, refine);
		} else {
			opaque_to_compiler(tx);
			o = encode_block_quad<CD, DXT, true, RAW, false>(tx, blk, width, 4, 4, refine);
		}
	} else if constexpr (LinW<CD>::value && REFINE != REFINE_LOOP) {
		if (DXT == DXT1 && __all((int)(RAW ? tx_min16(tx) >> 31 : tx_min16(tx) >> 24) != 0)) {
			opaque_to_compiler(tx);
			o = encode_block_lin<CD, DXT, true, RAW, true>(tx, blk, width, 4, 4, refine);
		} else {
			opaque_to_compiler(tx);
			o = encode_block_lin<CD, DXT, true, RAW, false>(tx, blk, width, 4, 4, refine);
		}
	} else {
		if (RAW) {
#pragma unroll
			for (int k = 0; k < 16; ++k) tx[k] = quant_none<DXT>(tx[k]);
		}
		o = encode_block_fast_t<CD, DXT, true>(tx, 4, 4, refine);
	}
	store_block<DXT>(dst, b, o);
}

// Waves per SIMD the one-image all-16 kernel is compiled for.  Its pair
// scans want more than 256 VGPRs; two waves (256 VGPRs, some values in
// scratch) hide memory latency better than one for every distance but the
// quadratic ones, whose DXT5 forms lose more to the scratch traffic than
// they gain.
template <int CD> struct All16Waves { static const int value = (CD == CD_AVG || CD == CD_WAVG || CD == CD_W0AVG) ? 1 : 2; };

// MODE_NORMAL without random colours over one image: one lane per block.
template <int CD, int DXT, bool RAW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(All16Waves<CD>::value))) void k_all16(const uint32_t *__restrict__ src, int width, int height,
                                               uint8_t *__restrict__ dst, int refine, int vec, int by0)
{
	S2TC_IMG_BLOCK_COORDS
	uint32_t tx[16];
	load_block<DXT, RAW>(src, width, bx, by, w, h, vec != 0, tx);
	const BlockOut o = encode_block_all16<CD, DXT>(tx, w, h, refine);
	store_block<DXT>(dst, b, o);
}

// MODE_NORMAL over one image: 16 lanes per block, 4 blocks per 64-lane
// workgroup; tile `tile` covers raster blocks 4*tile .. 4*tile+3.
template <int CD, bool RAW, bool DEDUP>
S2TC_D void normal_image_tile(const uint32_t *__restrict__ src, int width, int height, uint8_t *__restrict__ dst,
                              int dxt, int nrandom, int refine, unsigned long long seed, int pbw,
                              unsigned block_base, const GroupSmem<CD> &sm, int tile, int g, int t)
{
	const int bw = (width + 3) >> 2, bh = (height + 3) >> 2, nb = bw * bh;
	int b = tile * 4 + g;
	const bool have = b < nb;
	if (!have) b = nb - 1;
	const int by = b / bw, bx = b - by * bw;
	const int w = min(4, width - 4 * bx), h = min(4, height - 4 * by);
	const int x = t >> 2, y = t & 3;
	uint32_t tex = 0;
	if (x < w && y < h) {
		tex = src[(size_t)(4 * by + y) * (size_t)width + 4 * bx + x];
		if (RAW) tex = quant_none_rt(tex, dxt);
	}
	const uint32_t bkey = s2tc_prng_block_key((uint64_t)seed, block_base + (uint32_t)b);
	BlockOut o;
	encode_block_normal<CD, DEDUP>(tex, t, w, h, dxt, nrandom, refine, bkey, pbw, sm, o);
	if (have && t == 0) {
		if (dxt == DXT1) store_block<DXT1>(dst, (size_t)b, o);
		else store_block<DXT5>(dst, (size_t)b, o);
	}
}

// Candidates, rows and column features in LDS (nrandom <= S2TC_LDS_NRANDOM);
// one workgroup per tile.
template <int CD, bool RAW>
__global__ __launch_bounds__(64) void k_normal(const uint32_t *__restrict__ src, int width, int height,
                                               uint8_t *__restrict__ dst, int dxt, int nrandom, int refine,
                                               unsigned long long seed, int pbw, unsigned block_base)
{
	extern __shared__ __align__(16) uint8_t smem[];
	const int lane = (int)threadIdx.x;
	const int g = lane >> 4, t = lane & 15;
	const GroupSmem<CD> sm = group_smem<CD>(smem, g, 16 * pbw);
	normal_image_tile<CD, RAW, true>(src, width, height, dst, dxt, nrandom, refine, seed, pbw, block_base, sm,
	                                 xcd_local_id((int)blockIdx.x, (int)gridDim.x), g, t);
}

// More candidates than LDS holds: each workgroup keeps its four blocks'
// candidates, rows and column features in its own slice of `gmem`
// (4 * group_smem_bytes(16 * pbw) bytes per workgroup, gridDim.x slices) and
// strides over the tiles.  The slice is reused tile after tile, so a
// barrier ends every tile.
template <int CD, bool RAW>
__global__ __launch_bounds__(64) void k_normal_wide(const uint32_t *__restrict__ src, int width, int height,
                                                    uint8_t *__restrict__ dst, int dxt, int nrandom, int refine,
                                                    unsigned long long seed, int pbw, unsigned block_base,
                                                    uint8_t *__restrict__ gmem)
{
	const int lane = (int)threadIdx.x;
	const int g = lane >> 4, t = lane & 15;
	const size_t wg_bytes = 4 * group_smem_bytes<CD>(16 * pbw);
	const GroupSmem<CD> sm = group_smem<CD>(gmem + (size_t)blockIdx.x * wg_bytes, g, 16 * pbw);
	const int bw = (width + 3) >> 2, bh = (height + 3) >> 2;
	const int ntiles = (bw * bh + 3) >> 2;
	for (int tile = xcd_local_id((int)blockIdx.x, (int)gridDim.x); tile < ntiles; tile += (int)gridDim.x) {
		normal_image_tile<CD, RAW, false>(src, width, height, dst, dxt, nrandom, refine, seed, pbw, block_base, sm,
		                                  tile, g, t);
		__syncthreads();
	}
}

// ---- persistent batch kernels ----
//
// Jobs of one class (same distance; same format for quick mode) share a
// packed block index space: job j owns blocks [block0_j, block0_{j+1}).
// Tiles are TILE consecutive packed blocks; tile_jobs[t] = (first, last) job
// index the tile touches, so each lane or group finds its job by a short
// binary search.  Workgroups stride over the tiles (tile = blockIdx.x + k *
// gridDim.x), so the grid size is a free parameter and no counters need
// resetting between launches.

// by = b / bw, bx = b % bw for 0 <= b < 2^31 with bw and b / bw below 2^23
// (the batch plan rejects larger images), m = bw_magic(bw) = min(2^32 / bw,
// 2^32 - 1): b m / 2^32 lies in (b / bw - 1, b / bw], so its floor is the
// quotient or one less, and one correction step fixes it.  A quarter-rate
// high multiply and a few full-rate instructions instead of an integer
// division.
S2TC_D void div_blocks(int b, int bw, uint32_t m, int &by, int &bx)
{
	uint32_t q = __umulhi((uint32_t)b, m);
	uint32_t r = (uint32_t)b - (uint32_t)mul_i24((int)q, bw);
	const bool up = r >= (uint32_t)bw;
	q += up ? 1u : 0u;
	r -= up ? (uint32_t)bw : 0u;
	by = (int)q;
	bx = (int)r;
}

S2TC_D int find_job(const JobDesc *__restrict__ jobs, int lo, int hi, long long gb)
{
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (jobs[mid].block0 <= gb) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

template <int CD, int DXT, int REFINE>
__global__ __launch_bounds__(256) void k_fast_batch(const JobDesc *__restrict__ jobs, const TileJobs *__restrict__ tile_jobs,
                                                    long long nblocks, int ntiles)
{
	for (int tile = (int)blockIdx.x; tile < ntiles; tile += (int)gridDim.x) {
		const long long gb = (long long)tile * 256 + threadIdx.x;
		if (gb >= nblocks) continue;
		const TileJobs tj = tile_jobs[tile];
		const JobDesc &j = jobs[find_job(jobs, tj.lo, tj.hi, gb)];
		const int b = (int)(gb - j.block0);
		const int bw = (j.width + 3) >> 2;
		int by, bx;
		div_blocks(b, bw, j.bw_magic, by, bx);
		const int w = min(4, j.width - 4 * bx), h = min(4, j.height - 4 * by);
		const bool vec = (j.width & 3) == 0;
		uint32_t tx[16];
		BlockOut o;
		if constexpr (QuadW<CD>::value) {
			load_block<DXT, false>(j.src, j.width, bx, by, w, h, vec, tx);
			const uint32_t *blk = j.src + (size_t)(4 * by) * (size_t)j.width + 4 * bx;
			if (j.raw) o = encode_block_quad_any<CD, DXT, true>(tx, blk, j.width, w, h, REFINE);
			else o = encode_block_quad_any<CD, DXT, false>(tx, blk, j.width, w, h, REFINE);
		} else if constexpr (LinW<CD>::value && REFINE != REFINE_LOOP) {
			load_block<DXT, false>(j.src, j.width, bx, by, w, h, vec, tx);
			const uint32_t *blk = j.src + (size_t)(4 * by) * (size_t)j.width + 4 * bx;
			if (j.raw) o = encode_block_lin_any<CD, DXT, true>(tx, blk, j.width, w, h, REFINE);
			else o = encode_block_lin_any<CD, DXT, false>(tx, blk, j.width, w, h, REFINE);
		} else {
			if (j.raw) load_block<DXT, true>(j.src, j.width, bx, by, w, h, vec, tx);
			else load_block<DXT, false>(j.src, j.width, bx, by, w, h, vec, tx);
			o = encode_block_fast<CD, DXT>(tx, w, h, REFINE);
		}
		store_block<DXT>(j.dst, (size_t)b, o);
	}
}

template <int CD, int DXT, int REFINE>
S2TC_D void all16_batch_tiles(const JobDesc *__restrict__ jobs, const TileJobs *__restrict__ tile_jobs,
                              long long nblocks, int ntiles)
{
	for (int tile = (int)blockIdx.x; tile < ntiles; tile += (int)gridDim.x) {
		const long long gb = (long long)tile * 256 + threadIdx.x;
		if (gb >= nblocks) continue;
		const TileJobs tj = tile_jobs[tile];
		const JobDesc &j = jobs[find_job(jobs, tj.lo, tj.hi, gb)];
		const int b = (int)(gb - j.block0);
		const int bw = (j.width + 3) >> 2;
		int by, bx;
		div_blocks(b, bw, j.bw_magic, by, bx);
		const int w = min(4, j.width - 4 * bx), h = min(4, j.height - 4 * by);
		const bool vec = (j.width & 3) == 0;
		uint32_t tx[16];
		if (j.raw) load_block<DXT, true>(j.src, j.width, bx, by, w, h, vec, tx);
		else load_block<DXT, false>(j.src, j.width, bx, by, w, h, vec, tx);
		const BlockOut o = encode_block_all16<CD, DXT>(tx, w, h, REFINE);
		store_block<DXT>(j.dst, (size_t)b, o);
	}
}

// One wave per SIMD: inside a mixed batch the other classes fill the issue
// slots this kernel leaves idle, and at 256 VGPRs its DXT5 forms spill to
// scratch, whose write-backs multiplied its HBM writes tenfold for no gain.
template <int CD, int DXT, int REFINE>
__global__ __launch_bounds__(256) void k_all16_batch(const JobDesc *__restrict__ jobs,
                                                     const TileJobs *__restrict__ tile_jobs, long long nblocks,
                                                     int ntiles)
{
	all16_batch_tiles<CD, DXT, REFINE>(jobs, tile_jobs, nblocks, ntiles);
}

// The same for a class that has the GPU to itself (a batch of one kernel
// class, such as one large image): nothing else fills the idle issue slots,
// and two waves per SIMD are faster despite the scratch traffic.
template <int CD, int DXT, int REFINE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_all16_batch_alone(
    const JobDesc *__restrict__ jobs, const TileJobs *__restrict__ tile_jobs, long long nblocks, int ntiles)
{
	all16_batch_tiles<CD, DXT, REFINE>(jobs, tile_jobs, nblocks, ntiles);
}

template <int CD, bool DEDUP>
S2TC_D void normal_batch_tile(const JobDesc *__restrict__ jobs, const TileJobs *__restrict__ tile_jobs,
                              long long nblocks, const GroupSmem<CD> &sm, int tile, int g, int t)
{
	long long gb = (long long)tile * 4 + g;
	const bool have = gb < nblocks;
	if (!have) gb = nblocks - 1;
	const TileJobs tj = tile_jobs[tile];
	const JobDesc &j = jobs[find_job(jobs, tj.lo, tj.hi, gb)];
	const int b = (int)(gb - j.block0);
	const int bw = (j.width + 3) >> 2;
	int by, bx;
	div_blocks(b, bw, j.bw_magic, by, bx);
	const int w = min(4, j.width - 4 * bx), h = min(4, j.height - 4 * by);
	const int x = t >> 2, y = t & 3;
	uint32_t tex = 0;
	if (x < w && y < h) {
		tex = j.src[(size_t)(4 * by + y) * (size_t)j.width + 4 * bx + x];
		if (j.raw) tex = quant_none_rt(tex, j.dxt);
	}
	const uint32_t bkey = s2tc_prng_block_key((uint64_t)j.seed, (uint32_t)b);
	const int pbw = (16 + (j.nrandom > 0 ? j.nrandom : 0) + 15) / 16;
	BlockOut o;
	encode_block_normal<CD, DEDUP>(tex, t, w, h, j.dxt, j.nrandom, j.refine, bkey, pbw, sm, o);
	if (have && t == 0) {
		if (j.dxt == DXT1) store_block<DXT1>(j.dst, (size_t)b, o);
		else store_block<DXT5>(j.dst, (size_t)b, o);
	}
}

// One block's inputs, fetched a tile ahead of its encode so that the chain of
// dependent loads (tile -> job -> texels) overlaps the previous tile's work.
struct BlockIn {
	uint32_t tex;  // the texel word at a clamped address (every lane loads; see encode_fetched)
	int b, w, h, dxt, nrandom, refine, raw;
	uint32_t bkey;
	uint8_t *dst;
	bool have;
};

S2TC_D BlockIn fetch_block(const JobDesc *__restrict__ jobs, TileJobs tj, int tile, long long nblocks, int g, int t)
{
	BlockIn in;
	long long gb = (long long)tile * 4 + g;
	in.have = gb < nblocks;
	if (!in.have) gb = nblocks - 1;
	const JobDesc &j = jobs[find_job(jobs, tj.lo, tj.hi, gb)];
	in.b = (int)(gb - j.block0);
	const int bw = (j.width + 3) >> 2;
	int by, bx;
	div_blocks(in.b, bw, j.bw_magic, by, bx);
	in.w = min(4, j.width - 4 * bx);
	in.h = min(4, j.height - 4 * by);
	// an unconditional load (the address clamped into the block's valid
	// texels), so that nothing waits for it before the texel is used
	const int x = min(t >> 2, in.w - 1), y = min(t & 3, in.h - 1);
	in.tex = j.src[(size_t)(4 * by + y) * (size_t)j.width + 4 * bx + x];
	in.dxt = j.dxt;
	in.nrandom = j.nrandom;
	in.refine = j.refine;
	in.raw = j.raw;
	in.bkey = s2tc_prng_block_key((uint64_t)j.seed, (uint32_t)in.b);
	in.dst = j.dst;
	return in;
}

template <int CD> S2TC_D void encode_fetched(const BlockIn &in, const GroupSmem<CD> &sm, int t)
{
	const int x = t >> 2, y = t & 3;
	uint32_t tex = 0;
	if (x < in.w && y < in.h) tex = in.raw ? quant_none_rt(in.tex, in.dxt) : in.tex;
	const int pbw = (16 + (in.nrandom > 0 ? in.nrandom : 0) + 15) / 16;
	BlockOut o;
	encode_block_normal<CD, true>(tex, t, in.w, in.h, in.dxt, in.nrandom, in.refine, in.bkey, pbw, sm, o);
	if (in.have && t == 0) {
		if (in.dxt == DXT1) store_block<DXT1>(in.dst, (size_t)in.b, o);
		else store_block<DXT5>(in.dst, (size_t)in.b, o);
	}
}

// Software-pipelined over the workgroup's tiles: while tile T encodes, the
// job and texels of tile T + G are in flight and the tile record of T + 2G
// (G = grid size; indices past the end are clamped to the last tile, whose
// inputs are then fetched and never used).
template <int CD>
__global__ __launch_bounds__(64) void k_normal_batch(const JobDesc *__restrict__ jobs, const TileJobs *__restrict__ tile_jobs,
                                                     long long nblocks, int ntiles, int pbw_max)
{
	extern __shared__ __align__(16) uint8_t smem[];
	const int lane = (int)threadIdx.x;
	const int g = lane >> 4, t = lane & 15;
	const GroupSmem<CD> sm = group_smem<CD>(smem, g, 16 * pbw_max);
	const int G = (int)gridDim.x;
	int tile = xcd_local_id((int)blockIdx.x, G);
	if (tile >= ntiles) return;
	const auto tile_at = [&](int tl) {
		int i = min(tl, ntiles - 1);
		asm volatile("" : "+v"(i));  // a vector load: a scalar one would join the encoder's LDS waits
		return tile_jobs[i];
	};
	TileJobs tjn = tile_at(tile + G);
	BlockIn cur = fetch_block(jobs, tile_at(tile), tile, nblocks, g, t);
	for (; tile < ntiles; tile += G) {
		const BlockIn nxt = fetch_block(jobs, tjn, min(tile + G, ntiles - 1), nblocks, g, t);
		tjn = tile_at(tile + 2 * G);
		encode_fetched<CD>(cur, sm, t);
		__syncthreads();
		cur = nxt;
	}
}

// The batch form of k_normal_wide: group memory in a global slice per
// workgroup (4 * group_smem_bytes(16 * pbw_max) bytes each).
template <int CD>
__global__ __launch_bounds__(64) void k_normal_batch_wide(const JobDesc *__restrict__ jobs,
                                                          const TileJobs *__restrict__ tile_jobs, long long nblocks,
                                                          int ntiles, int pbw_max, uint8_t *__restrict__ gmem)
{
	const int lane = (int)threadIdx.x;
	const int g = lane >> 4, t = lane & 15;
	const size_t wg_bytes = 4 * group_smem_bytes<CD>(16 * pbw_max);
	const GroupSmem<CD> sm = group_smem<CD>(gmem + (size_t)blockIdx.x * wg_bytes, g, 16 * pbw_max);
	for (int tile = xcd_local_id((int)blockIdx.x, (int)gridDim.x); tile < ntiles; tile += (int)gridDim.x) {
		normal_batch_tile<CD, false>(jobs, tile_jobs, nblocks, sm, tile, g, t);
		__syncthreads();
	}
}

// Total colour distance of an encoding, one lane per block: the sum over
// the block's texels of ColorDist(texel, decoded colour) - the quantity
// s2tc_try_encode_block minimises (s2tc_algorithm.cpp:606-616) - with the
// colour decoded as fetch_2d_texel_* does (s2tc_libtxc_dxtn.cpp:35-140):
// index 0 / 1 -> c0 / c1, 2 and 3 -> checkerboard, DXT1 index 3 with
// c1 >= c0 -> black.  DXT1 texels with alpha 0 are skipped.  Each workgroup
// adds its int64 sum to *sum with one atomic.
template <int CD, int DXT, bool RAW>
__global__ __launch_bounds__(256) void k_colordist_error(const uint32_t *__restrict__ src, int width, int height,
                                                         const uint8_t *__restrict__ blocks, long long *sum)
{
	typedef Dist<CD> D;
	const int bw = (width + 3) >> 2, bh = (height + 3) >> 2;
	const int bx = (int)(blockIdx.x * 64 + threadIdx.x), by = (int)(blockIdx.y * 4 + threadIdx.y);
	long long e = 0;
	if (bx < bw && by < bh) {
		const int w = min(4, width - 4 * bx), h = min(4, height - 4 * by);
		const uint8_t *cb = blocks + ((size_t)by * (size_t)bw + (size_t)bx) * (DXT == DXT1 ? 8 : 16) + (DXT == DXT1 ? 0 : 8);
		const uint32_t c0 = cb[0] | (cb[1] << 8), c1 = cb[2] | (cb[3] << 8);
		const typename D::F f0 = D::feat(c0), f1 = D::feat(c1), fk = D::feat(0u);
		for (int y = 0; y < h; ++y) {
			const uint32_t row = cb[4 + y];
			for (int x = 0; x < w; ++x) {
				uint32_t t = src[(size_t)(4 * by + y) * (size_t)width + 4 * bx + x];
				if (RAW) t = quant_none<DXT>(t);
				if (DXT == DXT1 && texel_a(t) == 0) continue;
				const int b = (int)((row >> (2 * x)) & 3u);
				const bool odd = ((x ^ y) & 1) != 0;  // (i ^ j) & 1 with i, j multiples of 4 apart
				const typename D::F &fd = b == 0 ? f0 : b == 1 ? f1 : (DXT == DXT1 && b == 3 && c1 >= c0) ? fk : odd ? f1 : f0;
				e += D::d(D::featq(t), fd);
			}
		}
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
	__shared__ long long part[4];
	const int lane = (int)(threadIdx.y * 64 + threadIdx.x);
	if ((lane & 63) == 0) part[lane >> 6] = e;
	__syncthreads();
	if (lane == 0) atomicAdd((unsigned long long *)sum, (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

// Per-device tables of this kernel object's distance (one thread per RGB565
// colour, 256 x 256 threads).
template <int CD>
__global__ __launch_bounds__(256) void k_tables_init(int)
{
	const uint32_t v = blockIdx.x * 256u + threadIdx.x;
#ifdef S2TC_SRGB_Y_TABLE
	s2tc_srgb_y_table[v] = (uint16_t)srgb_get_y(c_r(v), c_g(v), c_b(v));
#endif
#ifdef S2TC_NM_TABLE
	const FeatF f = Dist<CD_NORMALMAP>::compute(v);
	s2tc_nm_table[v] = make_float4(f.x, f.y, f.z, 0.0f);
#endif
}

// Sets the cap of this kernel object's saturated pair-scan keys on the
// current device (one thread; the value is clamped to KEY_SAT_MAX).
template <int CD>
__global__ __launch_bounds__(64) void k_set_scan_cap(uint32_t cap)
{
	if (threadIdx.x == 0) s2tc_scan_cap = min(cap, KEY_SAT_MAX);
}

}  // namespace s2tc

using namespace s2tc;

#define S2TC_CAT2(a, b) a##b
#define S2TC_CAT(a, b) S2TC_CAT2(a, b)

extern "C" void S2TC_CAT(s2tc_kernels_cd, S2TC_CD)(KernelTable *kt)
{
#if S2TC_CD != 7  /* NORMALMAP never runs in quick mode (s2tc_algorithm.cpp:1131-1134) */
	kt->fast[DXT1][0] = (const void *)k_fast<S2TC_CD, DXT1, false>;
	kt->fast[DXT3][0] = (const void *)k_fast<S2TC_CD, DXT3, false>;
	kt->fast[DXT5][0] = (const void *)k_fast<S2TC_CD, DXT5, false>;
	kt->fast[DXT1][1] = (const void *)k_fast<S2TC_CD, DXT1, true>;
	kt->fast[DXT3][1] = (const void *)k_fast<S2TC_CD, DXT3, true>;
	kt->fast[DXT5][1] = (const void *)k_fast<S2TC_CD, DXT5, true>;
#define S2TC_INT3(D, R)                                                                      \
	kt->fast_int[D][R][0] = (const void *)k_fast_interior<S2TC_CD, D, R, REFINE_NEVER>;  \
	kt->fast_int[D][R][1] = (const void *)k_fast_interior<S2TC_CD, D, R, REFINE_ALWAYS>; \
	kt->fast_int[D][R][2] = (const void *)k_fast_interior<S2TC_CD, D, R, REFINE_LOOP>;
	S2TC_INT3(DXT1, false) S2TC_INT3(DXT3, false) S2TC_INT3(DXT5, false)
	S2TC_INT3(DXT1, true) S2TC_INT3(DXT3, true) S2TC_INT3(DXT5, true)
#undef S2TC_INT3
#else
	for (int d = 0; d < 3; ++d) {
		kt->fast[d][0] = kt->fast[d][1] = nullptr;
		for (int r = 0; r < 3; ++r) kt->fast_int[d][0][r] = kt->fast_int[d][1][r] = nullptr;
	}
#endif
#if S2TC_CD != 7
#define S2TC_FB3(D)                                                                \
	kt->fast_batch[D][0] = (const void *)k_fast_batch<S2TC_CD, D, REFINE_NEVER>;  \
	kt->fast_batch[D][1] = (const void *)k_fast_batch<S2TC_CD, D, REFINE_ALWAYS>; \
	kt->fast_batch[D][2] = (const void *)k_fast_batch<S2TC_CD, D, REFINE_LOOP>;
	S2TC_FB3(DXT1) S2TC_FB3(DXT3) S2TC_FB3(DXT5)
#undef S2TC_FB3
#else
	for (int d = 0; d < 3; ++d) kt->fast_batch[d][0] = kt->fast_batch[d][1] = kt->fast_batch[d][2] = nullptr;
#endif
	kt->normal_batch = (const void *)k_normal_batch<S2TC_CD>;
	kt->normal_batch_wide = (const void *)k_normal_batch_wide<S2TC_CD>;
	kt->error[DXT1][0] = (const void *)k_colordist_error<S2TC_CD, DXT1, false>;
	kt->error[DXT3][0] = (const void *)k_colordist_error<S2TC_CD, DXT3, false>;
	kt->error[DXT5][0] = (const void *)k_colordist_error<S2TC_CD, DXT5, false>;
	kt->error[DXT1][1] = (const void *)k_colordist_error<S2TC_CD, DXT1, true>;
	kt->error[DXT3][1] = (const void *)k_colordist_error<S2TC_CD, DXT3, true>;
	kt->error[DXT5][1] = (const void *)k_colordist_error<S2TC_CD, DXT5, true>;
	kt->all16[DXT1][0] = (const void *)k_all16<S2TC_CD, DXT1, false>;
	kt->all16[DXT3][0] = (const void *)k_all16<S2TC_CD, DXT3, false>;
	kt->all16[DXT5][0] = (const void *)k_all16<S2TC_CD, DXT5, false>;
	kt->all16[DXT1][1] = (const void *)k_all16<S2TC_CD, DXT1, true>;
	kt->all16[DXT3][1] = (const void *)k_all16<S2TC_CD, DXT3, true>;
	kt->all16[DXT5][1] = (const void *)k_all16<S2TC_CD, DXT5, true>;
#define S2TC_AB3(D)                                                                  \
	kt->all16_batch[D][0] = (const void *)k_all16_batch<S2TC_CD, D, REFINE_NEVER>;  \
	kt->all16_batch[D][1] = (const void *)k_all16_batch<S2TC_CD, D, REFINE_ALWAYS>; \
	kt->all16_batch[D][2] = (const void *)k_all16_batch<S2TC_CD, D, REFINE_LOOP>;
	S2TC_AB3(DXT1) S2TC_AB3(DXT3) S2TC_AB3(DXT5)
	for (int d = 0; d < 3; ++d)
		for (int r = 0; r < 3; ++r) kt->all16_batch_alone[d][r] = kt->all16_batch[d][r];
	if (All16Waves<S2TC_CD>::value == 2) {
#define S2TC_AL3(D)                                                                        \
	kt->all16_batch_alone[D][0] = (const void *)k_all16_batch_alone<S2TC_CD, D, REFINE_NEVER>;  \
	kt->all16_batch_alone[D][1] = (const void *)k_all16_batch_alone<S2TC_CD, D, REFINE_ALWAYS>; \
	kt->all16_batch_alone[D][2] = (const void *)k_all16_batch_alone<S2TC_CD, D, REFINE_LOOP>;
		S2TC_AL3(DXT1) S2TC_AL3(DXT3) S2TC_AL3(DXT5)
#undef S2TC_AL3
	}
#undef S2TC_AB3
	kt->normal[0] = (const void *)k_normal<S2TC_CD, false>;
	kt->normal[1] = (const void *)k_normal<S2TC_CD, true>;
	kt->normal_wide[0] = (const void *)k_normal_wide<S2TC_CD, false>;
	kt->normal_wide[1] = (const void *)k_normal_wide<S2TC_CD, true>;
#if defined(S2TC_NM_TABLE) || defined(S2TC_SRGB_Y_TABLE)
	kt->init = (const void *)k_tables_init<S2TC_CD>;
#else
	kt->init = nullptr;
#endif
	kt->set_scan_cap = (const void *)k_set_scan_cap<S2TC_CD>;
	kt->normal_smem_cand = (int)(4 * sizeof(uint32_t));
	kt->normal_smem_row = (int)(4 * S2TC_RS * sizeof(int));
	kt->normal_smem_fixed = (int)(4 * 16 * sizeof(Dist<S2TC_CD>::F));
}
