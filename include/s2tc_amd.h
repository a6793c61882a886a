/*
 * s2tc_amd.h - native (GPU) entry points of the MI355X S2TC encoder.
 *
 * These have no counterpart in the reference; they expose the GPU work at
 * the granularity it is efficient at (whole images, batches, a serving
 * loop).  The reference-compatible entry points are in txc_dxtn.h and
 * s2tc_algorithm.h.  All functions return 0 on success and a negative
 * S2TC_E* code on failure; s2tc_amd_last_error() gives the message.
 *
 * Enum values are those of s2tc_algorithm.h: DxtMode DXT1=0 DXT3=1 DXT5=2;
 * ColorDistMode RGB=0 YUV=1 SRGB=2 SRGB_MIXED=3 AVG=4 WAVG=5 W0AVG=6
 * NORMALMAP=7; RefinementMode NEVER=0 ALWAYS=1 LOOP=2; DitherMode NONE=0
 * SIMPLE=1 FLOYDSTEINBERG=2.
 *
 * Random candidates (nrandom > 0) come from the counter-based generator in
 * s2tc_prng.h keyed by (seed, raster block index within the image).
 */
#ifndef S2TC_AMD_H
#define S2TC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifndef S2TC_API
#if defined(__GNUC__)
#define S2TC_API __attribute__((visibility("default")))
#else
#define S2TC_API
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define S2TC_E_ARG      (-1)   /* invalid argument */
#define S2TC_E_HIP      (-2)   /* HIP runtime error */
#define S2TC_E_NOGPU    (-3)   /* no usable GPU */
#define S2TC_E_RANGE    (-4)   /* a size limit of the GPU path is exceeded */
#define S2TC_E_BUDGET   (-5)   /* request exceeds the serving memory budget */

/* Random candidates per block.  Up to S2TC_LDS_NRANDOM the candidate list
 * (16 + nrandom colours, rounded up to 16, 4 bytes each), its distance rows
 * from the 33rd on (64 bytes each; the first 32 stay in registers) and the
 * features of the first 16 candidates of four blocks fit the 64 KiB of LDS
 * a workgroup can allocate (224 candidates take 53,504 bytes).  Above it
 * the same encoder keeps them in device scratch memory (under m * 68 bytes
 * per block in flight, allocated stream-ordered per call, or with a batch).  S2TC_MAX_NRANDOM
 * bounds the candidate indices to 15 bits (16 + nrandom rounded up to 16
 * stays at or below 32768); larger values are S2TC_E_RANGE. */
#define S2TC_LDS_NRANDOM 208
#define S2TC_MAX_NRANDOM 32752

/* source kinds for s2tc_amd_encode_image */
#define S2TC_SRC_QUANTIZED 0   /* rgb565_image output: r5 g6 b5 a, 4 bytes per texel */
#define S2TC_SRC_RGBA8     1   /* 8-bit RGBA, quantized on load with DITHER_NONE */

S2TC_API const char *s2tc_amd_last_error(void);
S2TC_API int s2tc_amd_device_count(void);

/* Sets the calling thread's block counter: the generator key (with
 * S2TC_RANDOM_SEED) of the next per-block encoder call made through a
 * s2tc_encode_block_func pointer (s2tc_algorithm.h).  The counter starts at
 * 0 in every thread and advances by one per call. */
S2TC_API void s2tc_amd_set_block_index(unsigned int index);

/* device memory and stream plumbing (HIP device 0 unless set otherwise) */
S2TC_API int s2tc_amd_malloc(void **dptr, size_t bytes);
S2TC_API int s2tc_amd_free(void *dptr);
S2TC_API int s2tc_amd_memcpy_h2d(void *dst, const void *src, size_t bytes);
S2TC_API int s2tc_amd_memcpy_d2h(void *dst, const void *src, size_t bytes);
S2TC_API int s2tc_amd_memset(void *dst, int value, size_t bytes);
S2TC_API int s2tc_amd_synchronize(void);
/* Device-to-device copy by a plain streaming kernel that reads and writes
 * each byte once (16-byte aligned pointers and size); a known-size workload
 * for calibrating hardware byte counters. */
S2TC_API int s2tc_amd_memcpy_d2d_kernel(void *dst, const void *src, size_t bytes, void *stream);

/* Encode one image already resident on the device.  d_src: width*height
 * texels of 4 bytes (row stride = width); d_dst: ceil(w/4)*ceil(h/4) blocks
 * of 8 (DXT1) or 16 bytes in raster order.  Asynchronous on stream (a
 * hipStream_t, NULL = default stream). */
S2TC_API int s2tc_amd_encode_image(const void *d_src, int src_kind, int width, int height, void *d_dst,
                          int dxt, int cd, int nrandom, int refine, uint64_t seed, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, hosts that forbid
 * hipMalloc): s2tc_amd_encode_image with the same arguments and output,
 * calling no hipMalloc/hipFree or allocating routine.  Random-candidate
 * encodes above S2TC_LDS_NRANDOM take their group memory (under 4 * 68
 * bytes per candidate slot per workgroup) from a static device array of
 * S2TC_NOALLOC_SCRATCH bytes that each GPU holds until the process ends,
 * with as many workgroups as fit in it.  Calls that overlap on one GPU share
 * that array and race: serialize them (one stream, or one host thread). */
#define S2TC_NOALLOC_SCRATCH (64u << 20)
S2TC_API int s2tc_amd_encode_image_noalloc(const void *d_src, int src_kind, int width, int height, void *d_dst,
                                   int dxt, int cd, int nrandom, int refine, uint64_t seed, void *stream);

/* ---- batches: many images (textures x mip levels) in one graph launch ---- */

typedef struct {
	const void *src;   /* device: width*height texels, 4 bytes each (see src_kind) */
	void *dst;         /* device: ceil(width/4)*ceil(height/4) blocks in raster order */
	int width, height;
	int src_kind;      /* S2TC_SRC_QUANTIZED or S2TC_SRC_RGBA8 */
	int dxt, cd, nrandom, refine;
	uint64_t seed;     /* generator key of this image's random candidates */
} s2tc_amd_job;

/* per-kernel-class timing of a batch (s2tc_amd_batch_class_times) */
typedef struct {
	int normal;        /* 0: quick mode, 1: MODE_NORMAL with random colours (16 lanes
	                      per block), 2: MODE_NORMAL with the block colours only */
	int cd, dxt;       /* dxt is -1 for kind 1 classes (mixed formats) */
	int refine;        /* the refinement of all jobs of a kind 0 or 2 class, -1 for kind 1 */
	int njobs;
	long long blocks, texels;
	long long bytes;   /* algorithmic HBM bytes: 4 per texel read + block bytes written */
	double ms;         /* kernel time of one launch */
	int form;          /* kernel form (S2TC_FORM_*) the class runs */
} s2tc_amd_class_time;

/* Kernel forms of a batch class.  Quick and random-candidate classes have
 * one form (S2TC_FORM_LANE and S2TC_FORM_ROW16); classes of MODE_NORMAL
 * with the block colours only have up to three. */
#define S2TC_FORM_LANE        0   /* one lane per block, one wave per SIMD */
#define S2TC_FORM_LANE_2WAVES 1   /* one lane per block, two waves per SIMD (not for AVG, WAVG, W0AVG) */
#define S2TC_FORM_ROW16       2   /* 16 lanes (one DPP row) per block, four blocks per wave */

typedef struct s2tc_amd_batch s2tc_amd_batch;

/* Plans a batch: jobs are grouped into kernel classes (quick mode and
 * MODE_NORMAL without random colours per distance, format and refinement,
 * MODE_NORMAL with random colours per distance); each class runs one
 * persistent kernel over a packed list of 4x4 blocks, and all classes are
 * captured into one hipGraph.  The job array is copied; the device buffers
 * it points to must stay valid while the batch is used. */
S2TC_API int s2tc_amd_batch_create(const s2tc_amd_job *jobs, int njobs, s2tc_amd_batch **out);
/* Schedule: workgroups per CU of each persistent kernel (0 = one workgroup
 * per tile, i.e. not persistent) and the number of streams the classes are
 * spread over inside the graph.  Re-captures the graph. */
S2TC_API int s2tc_amd_batch_set_schedule(s2tc_amd_batch *b, int wg_per_cu, int streams);
/* Autotuner: times whole-graph launches (reps each) over a grid of
 * schedules, then over the kernel forms of each MODE_NORMAL class with the
 * block colours only (classes in decreasing cost, each form with the others
 * fixed), then over the schedule grid again, and keeps the fastest;
 * *best_ms gets its time per launch. */
S2TC_API int s2tc_amd_batch_autotune(s2tc_amd_batch *b, void *stream, int reps, double *best_ms);
/* Sets the kernel form of class `klass` (the order of
 * s2tc_amd_batch_class_times) and re-captures the graph; S2TC_E_ARG if the
 * class has no such form. */
S2TC_API int s2tc_amd_batch_set_form(s2tc_amd_batch *b, int klass, int form);
S2TC_API int s2tc_amd_batch_get_schedule(const s2tc_amd_batch *b, int *wg_per_cu, int *streams);
/* One launch of the whole batch (graph replay), asynchronous on stream. */
S2TC_API int s2tc_amd_batch_launch(s2tc_amd_batch *b, void *stream);
/* Launches each class alone, timed with HIP events; returns the class count
 * (at most max_out entries written). */
S2TC_API int s2tc_amd_batch_class_times(s2tc_amd_batch *b, void *stream, s2tc_amd_class_time *out, int max_out);
S2TC_API int s2tc_amd_batch_destroy(s2tc_amd_batch *b);

/* Decoding and encoding error (no counterpart among the reference's entry
 * points; for checking and measuring encodings on the device).
 *
 * s2tc_amd_decode_image: every texel of a block image with the rules of
 * fetch_2d_texel_{rgb_dxt1,rgba_dxt1,rgba_dxt3,rgba_dxt5}
 * (s2tc_libtxc_dxtn.cpp:35-140); format is the GL enum 0x83F0..0x83F3;
 * d_rgba receives width*height RGBA8 texels.  Asynchronous on `stream`.
 *
 * s2tc_amd_colordist_error: the sum over texels of ColorDist(texel, decoded
 * colour) in distance mode cd - the per-texel quantity the encoder minimises
 * (s2tc_algorithm.cpp:606-616) - for a source image (S2TC_SRC_QUANTIZED or
 * S2TC_SRC_RGBA8 quantized with DITHER_NONE) and its blocks; DXT1 texels with
 * alpha 0 are skipped.  Synchronous: *total is set on return. */
S2TC_API int s2tc_amd_decode_image(const void *d_blocks, unsigned int format, int width, int height, void *d_rgba,
                                   void *stream);
S2TC_API int s2tc_amd_colordist_error(const void *d_src, int src_kind, int width, int height, const void *d_blocks,
                                      int dxt, int cd, long long *total, void *stream);

/* The CLI's mip chain on the device: Image_MipReduce32
 * (s2tc_compress.c:427-493) halves an RGBA8 image with a truncating box
 * filter, repeated down to 1x1 (s2tc_compress.c:722-733).
 *
 * s2tc_amd_mip_layout: number of levels (level 0 = the image) and, for up to
 * max_levels of them, byte offsets in a chain buffer (each level 256-byte
 * aligned), sizes and the buffer's total size.
 * s2tc_amd_mip_reduce: one halving, d_in (width x height) -> d_out.
 * s2tc_amd_mip_chain: level 0 copied from d_src (no copy when d_src ==
 * d_levels: the image is already in place), then every level, into d_levels
 * laid out as s2tc_amd_mip_layout says; returns the level count.
 * Both are asynchronous on `stream`. */
S2TC_API int s2tc_amd_mip_layout(int width, int height, int *nlevels, size_t *offsets, int *widths, int *heights,
                                 int max_levels, size_t *total_bytes);
S2TC_API int s2tc_amd_mip_reduce(const void *d_in, int width, int height, void *d_out, void *stream);
S2TC_API int s2tc_amd_mip_chain(const void *d_src, int width, int height, void *d_levels, size_t capacity,
                                void *stream);

/* Many chains at once, one launch per level: each job's buffer holds level 0
 * in place and room for every level as s2tc_amd_mip_layout places them.
 * Returns after the chains are built (the stream is synchronized). */
typedef struct {
	void *levels;      /* device: the chain buffer, level 0 already filled */
	int width, height; /* level 0 */
} s2tc_amd_mip_job;
S2TC_API int s2tc_amd_mip_chain_batch(const s2tc_amd_mip_job *jobs, int njobs, void *stream);

/* s2tc_from_s3tc on the device: remaps nblocks S3TC blocks to S2TC
 * (s2tc_from_s3tc.cpp:77-190 per block, dispatched as :254-263).  dxt: 0
 * DXT1 (8-byte blocks, transparent mode kept), 1 DXT3, 2 DXT5 (16-byte
 * blocks).  d_in may equal d_out; buffers 8- (DXT1) or 16-byte aligned.
 * Asynchronous on `stream`. */
S2TC_API int s2tc_amd_from_s3tc(const void *d_in, void *d_out, size_t nblocks, int dxt, void *stream);

/* The same jobs with one plain kernel launch per job (the per-texture path). */
S2TC_API int s2tc_amd_encode_jobs_plain(const s2tc_amd_job *jobs, int njobs, void *stream);

/* ---- serving: requests arriving over time, continuous batching ---- */

typedef struct {
	double arrival_s;        /* arrival time in seconds after the start of the run (sorted) */
	int first_job, njobs;    /* the request's images: jobs[first_job .. first_job + njobs) */
} s2tc_amd_request;

typedef struct {
	size_t hbm_budget;       /* admitted, unfinished requests hold source + output bytes against it */
	int streams;             /* batches in flight at once (1..16), each on its own stream */
	int max_batch;           /* most requests in one dispatched batch */
	int serial_plain;        /* 1: one request at a time with one plain launch per image */
	int wg_per_cu;           /* persistent grid of the class kernels, rounded down to whole rounds of the
	                            workgroups a CU holds (0: one workgroup per tile); 16 serves the mixed
	                            pool best */
	void **keep_outputs;     /* NULL, or one host buffer per job receiving its blocks */
} s2tc_amd_serve_opts;

typedef struct {
	int completed, rejected; /* rejected: a request larger than the budget */
	int batches;
	int max_queue;           /* most requests waiting for admission at once */
	int max_inflight;        /* most requests admitted and unfinished at once */
	double wall_s;           /* start to last completion */
	long long texels;        /* texels of completed requests */
	double p50_ms, p99_ms, mean_ms, max_ms;  /* arrival-to-completion latency */
} s2tc_amd_serve_stats;

/* Runs the requests in real time: each arrives at its arrival_s, waits for
 * admission under the HBM budget, is batched with the other admitted
 * requests onto a free stream and completes when its batch has finished.
 * Job dst pointers are ignored: outputs live in a device arena owned by the
 * run (copied to keep_outputs when given).  latency_ms (nreq entries, may be
 * NULL) gets each request's latency, -1 for rejected ones. */
S2TC_API int s2tc_amd_serve(const s2tc_amd_job *jobs, int njobs, const s2tc_amd_request *reqs, int nreq,
                            const s2tc_amd_serve_opts *opts, double *latency_ms, s2tc_amd_serve_stats *stats);

/* ---- test hook ---- */

/* No counterpart in the reference; for tests only.  The pair scan of the
 * random-candidate encoder under RGB, YUV and SRGB_MIXED keys its pair sums
 * saturated at a cap (2^28 - 2), and a block whose least sum reaches the cap
 * scans again with exact 64-bit keys, so the output never depends on the
 * cap.  This sets the cap on the current device (clamped to 2^28 - 2, which
 * 0xffffffff restores); a low cap sends most blocks through the exact
 * rescan.  Call with no encode in flight. */
S2TC_API int s2tc_amd_debug_set_scan_cap(unsigned cap);

#ifdef __cplusplus
}
#endif

#endif
