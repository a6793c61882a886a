#!/usr/bin/env python3
"""Benchmark: S2TC encoding throughput on one MI355X.

Default workload ("batch"): 2000 RGBA8 textures of mixed power-of-two sizes
(4..2048 per side) with their full mip chains down to 1x1 - 16467 images,
7.54e8 texels - mixed DXT1/DXT3/DXT5 and all eight distance modes, encoded
bit-exactly (DITHER_NONE quantization) as one persistent hipGraph launch.
Sources and outputs are resident in HBM before the timed region.  A step is
one launch of the whole batch.

Printed (one JSON line): Mtexels/s of the batched launch (value), the same
jobs as plain per-image launches, the roofline of the dominant kernel class,
and the compiled reference's CPU throughput on a sample of the same jobs.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
                  [--workload batch|servepool|tiles4096|normal8192|sinus256]

servepool: the serving leg's pool of 256 textures (serving mix: 70% quick
mode, 30% nrandom 64 + REFINE_LOOP, full mip chains) as one batch - the
capacity that the serving loop schedules - checked against the reference's
per-texture digests.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
VALU_PEAK_TOPS = 39.3          # int32 lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz
# what the lane-op fractions count: the reference's pair scan for every
# block; the kernels skip it for blocks of one colour and scan repeated
# random colours once, so the fraction is a rate of the reference's work and
# can exceed 1 on smooth images
OPS_BASIS = ("n*m*(m-1) min+add lane-ops per block (x2 for DXT5 alpha) of the reference's "
             "reduce_colors_inplace, counted for every block")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per launch per kernel from the committed rocprofv3 counter
    summary (profiles/traffic_per_launch.txt): {kernel: (read, written)}."""
    out = {}
    path = os.path.join(ROOT, "profiles", "traffic_per_launch.txt")
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("#") or not line.strip():
                continue
            k, rd, wr = line.rstrip("\n").split("\t")
            out[k] = (int(rd), int(wr))
    return out


def image_items(s, name):
    """One single-image workload, quantized with its own dither (SIMPLE for
    all three, as tx_compress_dxtn does by default) before anything is timed;
    the items carry the quantized texels (SRC_QUANTIZED)."""
    from s2tc_amd import workloads
    wl = workloads.IMAGE_WORKLOADS[name]
    src = workloads.make_image_source(wl)
    q = s.rgb565_image(src, wl["width"], wl["height"], wl["srccomps"], {0: 1, 1: 4, 2: 8}[wl["dxt"]], wl["dither"])
    return [dict(rgba=q, width=wl["width"], height=wl["height"], dxt=wl["dxt"], cd=wl["cd"],
                 nrandom=wl["nrandom"], refine=wl["refine"], seed=wl["seed"])]


def cpu_model():
    """(CPU model name from /proc/cpuinfo, host cores this process may use)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, ncores


def median_of_5(fn):
    """One untimed warm-up call, then the median wall time of 5 calls."""
    import numpy as np
    fn()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _ref_job_type():
    class RefJob(ctypes.Structure):
        _fields_ = [("q", ctypes.c_void_p), ("dest", ctypes.c_void_p), ("width", ctypes.c_int),
                    ("height", ctypes.c_int), ("dxt", ctypes.c_int), ("cd", ctypes.c_int),
                    ("nrandom", ctypes.c_int), ("refine", ctypes.c_int), ("seed", ctypes.c_uint64)]
    return RefJob


def cpu_baseline(items, imgs=None, stride=5):
    """The compiled reference (oracle/_ref: the unmodified encoder sources),
    threads over all host cores, on every stride-th job of the workload (a
    fixed 1-in-`stride` sample: one pass over the full 16467-job list takes
    about 11 s on 16 cores, and 1 warm-up + 5 timed passes of it would keep
    the benchmark running for over a minute).  Quantization (DITHER_NONE, as the GPU
    path) is done before timing; the timed region is the reference's block
    loop only.  1 warm-up pass, median of 5."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    from oracle import Reference, reference_available
    if not reference_available():
        return None
    ref = Reference()
    RefJob = _ref_job_type()
    ref.lib.ref_encode_jobs_mt.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    ref.lib.ref_encode_jobs_mt.restype = ctypes.c_int
    model, ncores = cpu_model()
    keep, batch, texels = [], [], 0
    for i in range(0, len(items), stride):
        it = items[i]
        w, h = it["width"], it["height"]
        rgba = it["rgba"] if it["rgba"] is not None else imgs.source(i)
        q = ref.rgb565_image(rgba, w, h, 4, {0: 1, 1: 4, 2: 8}[it["dxt"]], 0)
        dst = np.zeros(((w + 3) // 4) * ((h + 3) // 4) * (8 if it["dxt"] == 0 else 16), np.uint8)
        keep += [q, dst]
        batch.append(RefJob(q.ctypes.data, dst.ctypes.data, w, h, it["dxt"], it["cd"], it["nrandom"],
                            it["refine"], it["seed"]))
        texels += w * h
    ja = (RefJob * len(batch))(*batch)
    used = [0]

    def run():
        used[0] = ref.lib.ref_encode_jobs_mt(ctypes.cast(ja, ctypes.c_void_p), len(batch), ncores)
    t = median_of_5(run)
    return dict(value=texels / t / 1e6, unit="Mtexels/s", cores=int(used[0]), kind="reference", cpu_model=model,
                sample=f"every {stride}th job of the same workload: {len(batch)} of {len(items)} images, "
                       f"{texels} texels; DITHER_NONE quantization before timing, the reference's block loop "
                       f"timed, {used[0]} threads, 1 warm-up + median of 5 passes ({t:.2f} s each)")


def image_configs_leg(s, with_cpu=True, reps=5):
    """The two single-image configurations, device-resident, next to the
    compiled reference on the host cores:
      tiles4096  4096x4096 RGBA8 tiles, DXT5, SRGB_MIXED, all 16 candidates, LOOP
      normal8192 8192x8192 RGB8 normal map, DXT1, NORMALMAP, nrandom 64, LOOP
    Both quantized with the library default SIMPLE dither first (as
    tx_compress_dxtn does; not timed).  GPU: encode kernels on the quantized
    image in HBM, HIP events, median of `reps` launches after one warm-up.
    CPU: the reference's block loop over a band of the same quantized image
    (the top 1024 block rows' texels of tiles4096, 64 texel rows of
    normal8192: the whole images would take 2 s and 93 s per pass), all host
    cores, 1 warm-up + median of 5."""
    import numpy as np
    from s2tc_amd import workloads
    ref = None
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Reference, reference_available
        ref = Reference() if reference_available() else None
    model, ncores = cpu_model()
    ev = HipEvents()
    out = []
    for name, band in (("tiles4096", 1024), ("normal8192", 64)):
        wl = workloads.IMAGE_WORKLOADS[name]
        w, h = wl["width"], wl["height"]
        src = workloads.make_image_source(wl)
        q = s.rgb565_image(src, w, h, wl["srccomps"], {0: 1, 1: 4, 2: 8}[wl["dxt"]], wl["dither"])
        del src
        d_q = s.DeviceBuffer.from_host(q)
        nb = s.nblocks(w, h)
        d_o = s.DeviceBuffer(nb * s.block_bytes(wl["dxt"]))
        args = (d_q, s.SRC_QUANTIZED, w, h, d_o, wl["dxt"], wl["cd"], wl["nrandom"], wl["refine"], wl["seed"])
        s.encode_image_device(*args)
        s.synchronize()
        ts = []
        for _ in range(reps):
            ev.record(0)
            s.encode_image_device(*args)
            ev.record(1)
            ts.append(ev.elapsed_ms())
        ms = float(np.median(ts))
        m = 16 + max(0, wl["nrandom"])
        scans = nb
        if wl["dxt"] == 2:
            # the DXT5 alpha scan runs for every block but the full ones with one alpha value
            fw, fh = w // 4, h // 4
            al = q.reshape(h, w, 4)[:4 * fh, :4 * fw, 3].reshape(fh, 4, fw, 4)
            scans += nb - int((al == al[:, :1, :, :1]).all(axis=(1, 3)).sum())
        ops = 16 * m * (m - 1) * scans
        row = dict(workload=name, texels=w * h, kernel_ms=round(ms, 4), mtexels_s=round(w * h / (ms * 1e-3) / 1e6, 1),
                   valu_floor_ops=ops, valu_frac=round(ops / (ms * 1e-3) / 1e12 / VALU_PEAK_TOPS, 4),
                   ops_basis=OPS_BASIS)
        if ref is not None:
            bh = min(h, band)
            qb = np.ascontiguousarray(q[:w * bh * 4])
            used = [0]

            def run():
                used[0] = ref.encode_quantized_mt(qb, w, bh, wl["dxt"], wl["cd"], wl["nrandom"], wl["refine"],
                                                  seed=wl["seed"], nthreads=ncores)[1]
            t = median_of_5(run)
            row["cpu_reference"] = dict(value=round(w * bh / t / 1e6, 2), unit="Mtexels/s", cores=int(used[0]),
                                        cpu_model=model, sample=f"top {bh} of {h} texel rows ({w * bh} texels), "
                                                                f"1 warm-up + median of 5")
        log(f"config {row}")
        out.append(row)
        d_q.free()
        d_o.free()
    ev.close()
    return out


def alpha_scan_blocks(imgs, cd, refine):
    """Blocks of the all-candidate DXT5 class (cd, refine) that run the alpha
    pair scan: all but the full blocks whose 16 alphas are equal."""
    import numpy as np
    n = 0
    for i, it in enumerate(imgs.items):
        fast = it["nrandom"] < 0 and it["cd"] != 7
        if it["dxt"] != 2 or it["cd"] != cd or it["refine"] != refine or fast or it["nrandom"] > 0:
            continue
        w, h = it["width"], it["height"]
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        fw, fh = w // 4, h // 4          # the full blocks; partial ones always scan
        uni = 0
        if fw and fh:
            a = imgs.source(i).reshape(h, w, 4)[:4 * fh, :4 * fw, 3].reshape(fh, 4, fw, 4)
            uni = int((a == a[:, :1, :, :1]).all(axis=(1, 3)).sum())
        n += nb - uni
    return n


class HipEvents:
    """HIP events on the null stream (the stream encode_image_device launches on)."""
    def __init__(self):
        try:
            self.h = ctypes.CDLL("libamdhip64.so")
        except OSError:
            self.h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.ev = [ctypes.c_void_p(), ctypes.c_void_p()]
        for e in self.ev:
            assert self.h.hipEventCreate(ctypes.byref(e)) == 0

    def record(self, i):
        assert self.h.hipEventRecord(self.ev[i], None) == 0

    def elapsed_ms(self):
        assert self.h.hipEventSynchronize(self.ev[1]) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(ms), self.ev[0], self.ev[1]) == 0
        return ms.value

    def close(self):
        for e in self.ev:
            self.h.hipEventDestroy(e)


def quick_leg(s, size=16384, reps=10, cds=(0, 1, 2, 3, 4, 5, 6)):
    """Quick mode (MODE_FAST) on large RGBA8 images in HBM, every distance
    that has a quick mode (all but NORMALMAP), REFINE_ALWAYS, DITHER_NONE
    quantization on load; k_fast_interior launches timed with HIP events.
    DXT1 on the sinusoid + noise content of the 256x256 case at size x size
    (opaque), DXT5 on the gradient-tiles content (radial alpha with 0/255
    speckle: the alpha endpoint search runs on almost every block).
    Algorithmic bytes: 4 per texel read + 8 (DXT1) or 16 (DXT5) per block
    written."""
    from s2tc_amd import synth
    names = s.COLORDIST_NAMES
    ev = HipEvents()
    out = []
    traffic = measured_traffic()
    for dxt, gen in ((0, synth.sinus), (2, synth.tiles)):
        src = gen(size, size, 0)
        d_src = s.DeviceBuffer.from_host(src)
        del src
        nb = s.nblocks(size, size)
        d_dst = s.DeviceBuffer(nb * s.block_bytes(dxt))
        for cd in cds:
            s.encode_image_device(d_src, s.SRC_RGBA8, size, size, d_dst, dxt, cd, -1, s.REFINE_ALWAYS)
            s.synchronize()
            ev.record(0)
            for _ in range(reps):
                s.encode_image_device(d_src, s.SRC_RGBA8, size, size, d_dst, dxt, cd, -1, s.REFINE_ALWAYS)
            ev.record(1)
            ms = ev.elapsed_ms() / reps
            texels = size * size
            nbytes = 4 * texels + nb * s.block_bytes(dxt)
            gbs = nbytes / (ms * 1e-3) / 1e9
            kname = f"k_fast_interior<{cd}, {dxt}, true, 1>"
            tr = traffic.get(kname) if size == 16384 else None
            out.append(dict(kernel=kname, cd=names[cd], image=f"{size}x{size} {gen.__name__}", dxt=(1, 3, 5)[dxt],
                            texels=texels, kernel_ms=round(ms, 4), mtexels_s=round(texels / (ms * 1e-3) / 1e6, 1),
                            algorithmic_bytes=nbytes, achieved_gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4),
                            traffic=sum(tr) if tr else None))
            log(f"quick {out[-1]}")
        d_dst.free()
        if dxt == 0:
            d_src.free()
    # a plain streaming copy of the same source, for the achievable read+write rate
    d_cp = s.DeviceBuffer(d_src.nbytes)
    L = s.lib()
    L.s2tc_amd_memcpy_d2d_kernel.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    s.check(L.s2tc_amd_memcpy_d2d_kernel(d_cp.ptr, d_src.ptr, d_src.nbytes, None))
    s.synchronize()
    ev.record(0)
    for _ in range(reps):
        s.check(L.s2tc_amd_memcpy_d2d_kernel(d_cp.ptr, d_src.ptr, d_src.nbytes, None))
    ev.record(1)
    ms = ev.elapsed_ms() / reps
    out.append(dict(kernel="k_copy16 (reference copy)", bytes=2 * d_src.nbytes, kernel_ms=round(ms, 4),
                    achieved_gbs=round(2 * d_src.nbytes / (ms * 1e-3) / 1e9, 1)))
    log(f"quick {out[-1]}")
    d_cp.free()
    ev.close()
    d_src.free()
    return out


def block_call_row(s, ref, reps=200):
    """The per-block entry points of s2tc_encode_block_func (one 4x4 block of
    quantized texels per call, synchronous: one upload, launch and download
    each), median over `reps` calls, beside the compiled reference's own
    per-block function on the same block."""
    import numpy as np
    FT = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int)
    rng = np.random.default_rng(3)
    q = np.stack([rng.integers(0, 32, 16), rng.integers(0, 64, 16), rng.integers(0, 32, 16),
                  rng.integers(0, 256, 16)], -1).astype(np.uint8).reshape(-1)
    out = np.zeros(16, np.uint8)

    def time_lib(lib):
        lib.s2tc_encode_block_func.restype = ctypes.c_void_p
        lib.s2tc_encode_block_func.argtypes = [ctypes.c_int] * 4
        f = FT(lib.s2tc_encode_block_func(2, 5, -1, 1))  # DXT5, WAVG, quick selection, REFINE_ALWAYS
        for _ in range(10):
            f(out.ctypes.data, q.ctypes.data, 4, 4, 4, -1)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f(out.ctypes.data, q.ctypes.data, 4, 4, 4, -1)
            ts.append(time.perf_counter() - t0)
        return round(float(np.median(ts)) * 1e6, 2)
    row = dict(workload="one 4x4 block, DXT5, WAVG, quick, REFINE_ALWAYS", entry="s2tc_encode_block_func",
               texels=16, us_per_call=time_lib(s.lib()), calls=reps)
    if ref is not None:
        row["cpu_reference_us_per_call"] = time_lib(ref.lib)
    log(f"calls {row}")
    return row


def calls_leg(s, with_cpu=True, which=("sinus256", "tiles4096", "normal8192")):
    """The drop-in entry point tx_compress_dxtn on the single-image workloads,
    timed end to end per call (host buffers in and out: upload, SIMPLE dither,
    encode, download; synchronous), median of several calls.  The 256x256
    case is the reference CLI's per-level call, launch-latency-bound; beside
    it the compiled reference's own tx_compress_dxtn (one host thread, the
    reference being single-threaded per call) on the same image."""
    import numpy as np
    from s2tc_amd import workloads
    names = ["RGB", "YUV", "SRGB", "SRGB_MIXED", "AVG", "WAVG", "W0AVG", "NORMALMAP"]
    L = s.lib()
    L.tx_compress_dxtn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint,
                                   ctypes.c_void_p, ctypes.c_int]
    ref = None
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Reference, reference_available
        ref = Reference() if reference_available() else None
    saved = {k: os.environ.get(k) for k in ("S2TC_DITHER_MODE", "S2TC_COLORDIST_MODE", "S2TC_RANDOM_COLORS",
                                            "S2TC_REFINE_COLORS", "S2TC_RANDOM_SEED")}
    out = []
    try:
        for name, reps in (("sinus256", 50), ("tiles4096", 5), ("normal8192", 3)):
            if name not in which:
                continue
            wl = workloads.IMAGE_WORKLOADS[name]
            src = workloads.make_image_source(wl)
            w, h = wl["width"], wl["height"]
            os.environ.update(S2TC_DITHER_MODE=["NONE", "SIMPLE", "FLOYDSTEINBERG"][wl["dither"]],
                              S2TC_COLORDIST_MODE=names[wl["cd"]], S2TC_RANDOM_COLORS=str(wl["nrandom"]),
                              S2TC_REFINE_COLORS=["NEVER", "ALWAYS", "LOOP"][wl["refine"]],
                              S2TC_RANDOM_SEED=str(wl["seed"]))
            bs = 8 if wl["dxt"] == 0 else 16
            bw, bh = (w + 3) // 4, (h + 3) // 4
            dst = np.zeros(bw * bh * bs, np.uint8)
            fmt = {0: 0x83F1, 1: 0x83F2, 2: 0x83F3}[wl["dxt"]]

            def call(lib):
                lib.tx_compress_dxtn(wl["srccomps"], w, h, src.ctypes.data, fmt, dst.ctypes.data, bw * bs)
            call(L)
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                call(L)
                ts.append(time.perf_counter() - t0)
            t = float(np.median(ts))
            row = dict(workload=name, entry="tx_compress_dxtn", texels=w * h, us_per_call=round(t * 1e6, 1),
                       mtexels_s_end_to_end=round(w * h / t / 1e6, 1), calls=reps)
            if ref is not None and name == "sinus256":
                call(ref.lib)
                rs = []
                for _ in range(10):
                    t0 = time.perf_counter()
                    call(ref.lib)
                    rs.append(time.perf_counter() - t0)
                row["cpu_reference_us_per_call"] = round(float(np.median(rs)) * 1e6, 1)
            log(f"calls {row}")
            out.append(row)
        out.append(block_call_row(s, ref))
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return out


def side_leg(s, reps=10):
    """The kernels beside the encoder, each HBM-bound, timed with HIP events:
    one mip halving of a 16384x16384 RGBA8 image (k_mip_reduce: 1 GiB read,
    256 MiB written) and the S3TC -> S2TC remap of 2^24 DXT1 and DXT5 blocks
    (k_from_s3tc: every block read and written once)."""
    import numpy as np
    ev = HipEvents()
    out = []
    rng = np.random.default_rng(9)
    n = 16384
    d_src = s.DeviceBuffer.from_host(rng.integers(0, 256, n * n * 4, dtype=np.uint8))
    d_dst = s.DeviceBuffer(n * n)
    s.mip_reduce_device(d_src, n, n, d_dst)
    s.synchronize()
    ev.record(0)
    for _ in range(reps):
        s.mip_reduce_device(d_src, n, n, d_dst)
    ev.record(1)
    ms = ev.elapsed_ms() / reps
    nbytes = n * n * 4 + n * n
    out.append(dict(kernel="k_mip_reduce", image=f"{n}x{n} -> {n // 2}x{n // 2}", kernel_ms=round(ms, 4),
                    algorithmic_bytes=nbytes, achieved_gbs=round(nbytes / (ms * 1e-3) / 1e9, 1),
                    frac=round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)))
    log(f"side {out[-1]}")
    d_dst.free()
    for dxt in (0, 2):
        nb = 1 << 24
        bs = s.block_bytes(dxt)
        d_out = s.DeviceBuffer(nb * bs)
        s.from_s3tc_device(d_src, d_out, nb, dxt)
        s.synchronize()
        ev.record(0)
        for _ in range(reps):
            s.from_s3tc_device(d_src, d_out, nb, dxt)
        ev.record(1)
        ms = ev.elapsed_ms() / reps
        nbytes = 2 * nb * bs
        out.append(dict(kernel=f"k_from_s3tc<{dxt}>", blocks=nb, dxt=(1, 3, 5)[dxt], kernel_ms=round(ms, 4),
                        algorithmic_bytes=nbytes, achieved_gbs=round(nbytes / (ms * 1e-3) / 1e9, 1),
                        frac=round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)))
        log(f"side {out[-1]}")
        d_out.free()
    d_src.free()
    ev.close()
    return out


def quality_leg(s, w=1024, h=512):
    """Total colour distance (s2tc_amd_colordist_error) of a NORMALMAP,
    nrandom 64, REFINE_LOOP, DXT1 encoding of a w x h normal map (512
    Gaussians, as the 8192x8192 case), with this encoder's counter-based
    generator and with the unmodified reference drawing from libc's rand()
    (oracle/_ref/libs2tc_ref_stock.so, single thread, srand(1))."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import StockReference, stock_reference_available
    if not stock_reference_available():
        return None
    import numpy as np
    from s2tc_amd import synth
    rgba = synth.normalmap(w, h, 2, 512, 4)
    q = rgba.reshape(-1, 4).copy()
    q[:, 0] >>= 3
    q[:, 1] >>= 2
    q[:, 2] >>= 3
    q[:, 3] >>= 7
    ours = s.encode_image(rgba, w, h, 0, s.NORMALMAP, 64, s.REFINE_LOOP, seed=2, src_kind=s.SRC_RGBA8)
    e_ours = s.colordist_error(rgba, w, h, ours, 0, s.NORMALMAP, src_kind=s.SRC_RGBA8)
    st = StockReference()
    st.srand(1)
    t0 = time.perf_counter()
    stock = st.encode_quantized(q.reshape(-1), w, h, 0, s.NORMALMAP, 64, s.REFINE_LOOP)
    t_stock = time.perf_counter() - t0
    e_stock = s.colordist_error(rgba, w, h, stock, 0, s.NORMALMAP, src_kind=s.SRC_RGBA8)
    out = dict(workload=f"{w}x{h} normal map, NORMALMAP, nrandom 64, LOOP, DXT1", colordist_error=e_ours,
               stock_reference_colordist_error=e_stock, ratio=round(e_ours / max(e_stock, 1), 5),
               stock_reference_seconds=round(t_stock, 2))
    log(f"quality {out}")
    return out


def serve_leg(s, rates, duration, npool=256, seed=4):
    """Open-loop Poisson arrivals of textures (with mip chains) drawn from a
    pool of the serving mix (70% quick mode, 30% nrandom 64 + LOOP), sources
    resident on the device, 64 GiB admission budget; continuous batching on 4
    streams vs one-at-a-time plain launches."""
    import numpy as np
    from s2tc_amd import workloads
    textures = workloads.batch_textures(seed, npool, serve_mix=True)
    imgs = s.DeviceImages.from_textures(textures)
    groups, f = [], 0
    for t in textures:
        n = max(t["width"], t["height"]).bit_length()   # levels down to 1x1
        groups.append((f, n))
        f += n
    assert f == len(imgs.items)
    del textures
    out = {}
    for mode in ("batched", "serial_plain"):
        for rate in rates:
            if mode == "serial_plain" and rate > 10000:
                continue
            rng = np.random.default_rng(rate)
            n = max(1, int(rate * duration))
            arr = np.cumsum(rng.exponential(1.0 / rate, n))
            pick = rng.integers(0, npool, n)
            reqs = [(float(a), groups[p][0], groups[p][1]) for a, p in zip(arr, pick)]
            st, lat = s.serve(imgs, reqs, hbm_budget=64 << 30, streams=4, max_batch=256,
                              serial_plain=(mode == "serial_plain"))
            out[f"{mode}@{rate}"] = dict(requests=n, p50_ms=round(st["p50_ms"], 3), p99_ms=round(st["p99_ms"], 3),
                                        mean_ms=round(st["mean_ms"], 3), max_ms=round(st["max_ms"], 3),
                                        mtexels_s=round(st["texels"] / max(st["wall_s"], 1e-9) / 1e6, 1),
                                        batches=st["batches"], max_queue=st["max_queue"],
                                        offered_mtexels_s=round(st["texels"] / arr[-1] / 1e6, 1))
            log(f"serve {mode}@{rate}: {out[f'{mode}@{rate}']}")
    imgs.free()
    return out


def verify_batch_output(imgs, workload):
    """SHA-256 over the per-image block digests of the batch output (batch),
    or per texture over its mip levels' blocks (servepool), against the
    compiled reference's digests of the same jobs (tests/golden/
    digests.json).  True / False, or None where no digest exists."""
    import hashlib
    digests = json.load(open(os.path.join(ROOT, "tests", "golden", "digests.json")))["workloads"]
    exp = digests.get({"batch": "batch", "servepool": "serve_pool"}.get(workload, ""))
    if not exp or len(imgs.items) != exp["jobs"]:
        return None
    host = imgs.dst.to_host()
    if workload == "servepool":
        f, ok = 0, True
        for n, want in zip(imgs.levels_per_texture, exp["texture_sha256"]):
            got = hashlib.sha256(b"".join(imgs.output(f + q, host).tobytes() for q in range(n))).hexdigest()
            ok = ok and got == want
            f += n
        return ok and f == len(imgs.items)
    hashes = b"".join(hashlib.sha256(imgs.output(i, host).tobytes()).digest() for i in range(len(imgs.items)))
    return hashlib.sha256(hashes).hexdigest() == exp["sha256_of_job_sha256s"]


def class_roofline(s, c, imgs, nclasses):
    """Adds to a class_times entry: its kernel's name, HBM rate and fraction
    (algorithmic bytes: 4 per texel read + the block bytes written) and, for
    the all-16 and random-candidate classes, the pair-scan floor of
    reduce_colors_inplace - n m (m - 1) min+add int32 lane-ops per block with
    n = 16 texel colours and m = 16 + nrandom candidates, plus the same again
    for the DXT5 alpha scan of every block that runs one (full blocks with a
    single alpha value run none) - and its fraction of the VALU peak."""
    # all-16 classes run the kernel form the autotuner chose: one lane per
    # block with one (k_all16_batch) or two (k_all16_batch_alone) waves per
    # SIMD, or 16 lanes per block (k_normal_batch)
    all16 = {0: f"k_all16_batch<{c['cd']}, {c['dxt']}, {c['refine']}>",
             1: f"k_all16_batch_alone<{c['cd']}, {c['dxt']}, {c['refine']}>",
             2: f"k_normal_batch<{c['cd']}>"}[c["form"]]
    c["kernel_name"] = {0: f"k_fast_batch<{c['cd']}, {c['dxt']}, {c['refine']}>", 1: f"k_normal_batch<{c['cd']}>",
                        2: all16}[c["normal"]]
    c["gbs"] = c["bytes"] / (c["ms"] * 1e-3) / 1e9
    c["hbm_frac"] = c["gbs"] / HBM_PEAK_GBS
    ops = 0
    if c["normal"] == 2:
        scans = c["blocks"]
        if c["dxt"] == 2:
            scans += alpha_scan_blocks(imgs, c["cd"], c["refine"])
        ops = 16 * 16 * 15 * scans
    elif c["normal"] == 1:
        for it in imgs.items:
            if it["cd"] == c["cd"] and it["nrandom"] > 0:
                m = 16 + it["nrandom"]
                ops += s.nblocks(it["width"], it["height"]) * 16 * m * (m - 1) * (2 if it["dxt"] == 2 else 1)
    c["valu_ops"] = ops
    c["valu_tops"] = ops / (c["ms"] * 1e-3) / 1e12
    c["valu_frac"] = c["valu_tops"] / VALU_PEAK_TOPS


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batch")
    ap.add_argument("--schedule", default="", help="wg_per_cu,streams (skips the autotuner)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-quick", action="store_true")
    ap.add_argument("--quick-size", type=int, default=16384)
    ap.add_argument("--no-side", action="store_true")
    ap.add_argument("--no-calls", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--profile-classes", action="store_true",
                    help="only launch each kernel class alone, --steps times (give --schedule to skip the "
                         "autotuner's graph launches), and print the class times")
    ap.add_argument("--only-calls", action="store_true",
                    help="only time tx_compress_dxtn per call on the single-image workloads and print their JSON")
    ap.add_argument("--calls-workloads", default="sinus256,tiles4096,normal8192")
    ap.add_argument("--only-side", action="store_true",
                    help="only run the mip / transcode kernels and print their JSON")
    ap.add_argument("--only-quick", action="store_true",
                    help="only run the quick-mode leg (one large image) and print its JSON")
    ap.add_argument("--serve-duration", type=float, default=0.5)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a plain device copy of this many MiB (PMC calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for {a.gpus}); textures shard across GPUs by "
              "whole job with no communication, so N GPUs are N independent copies of this run.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import s2tc_amd as s
    L = s.lib()
    assert L.s2tc_amd_device_count() > 0, "no GPU visible"
    if a.calibrate_copy:
        n = a.calibrate_copy << 20
        src, dst = s.DeviceBuffer(n), s.DeviceBuffer(n)
        src.fill(1)
        L.s2tc_amd_memcpy_d2d_kernel.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        for _ in range(3):
            s.check(L.s2tc_amd_memcpy_d2d_kernel(dst.ptr, src.ptr, n, None))
        s.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n}))
        return

    if a.only_calls:
        print(json.dumps({"calls": calls_leg(s, not a.no_cpu, a.calls_workloads.split(","))}), flush=True)
        return

    if a.only_side:
        print(json.dumps({"side_kernels": side_leg(s, max(1, a.steps))}), flush=True)
        return

    if a.only_quick:
        print(json.dumps({"quick_mode": quick_leg(s, a.quick_size, max(1, a.steps))}), flush=True)
        return

    t0 = time.time()
    mip_leg = None
    if a.workload in ("batch", "servepool"):
        from s2tc_amd import workloads
        textures = workloads.batch_textures(3, 2000) if a.workload == "batch" else \
            workloads.batch_textures(workloads.SERVE_SEED, workloads.SERVE_POOL, serve_mix=True)
        log(f"workload {a.workload}: {len(textures)} base levels generated in {time.time() - t0:.1f} s")
        t1 = time.perf_counter()
        imgs = s.DeviceImages.from_textures(textures)
        mip_leg = dict(textures=len(textures), uploaded_bytes=imgs.uploaded_bytes,
                       upload_and_chain_s=round(time.perf_counter() - t1, 3))
        imgs.levels_per_texture = [max(t["width"], t["height"]).bit_length() for t in textures]
        del textures
        items = imgs.items
    else:
        items = image_items(s, a.workload)
        imgs = s.DeviceImages(items, src_kind=s.SRC_QUANTIZED)
    log(f"workload {a.workload}: {len(items)} images on the device after {time.time() - t0:.1f} s")
    texels = imgs.texels
    batch = s.Batch(imgs)
    if a.schedule:
        wg, ns = (int(v) for v in a.schedule.split(","))
        batch.set_schedule(wg, ns)
        tuned_ms = None
    else:
        tuned_ms = batch.autotune(reps=2)
    sched = batch.schedule()
    log(f"schedule wg_per_cu={sched[0]} streams={sched[1]} (autotuned {tuned_ms} ms)")

    if a.profile_classes:
        # standalone launches of each class only (no graph replays), so that a
        # kernel trace's per-kernel averages are the class times printed here
        for _ in range(a.steps):
            ct = batch.class_times()
        for c in ct:
            class_roofline(s, c, imgs, len(ct))
        print(json.dumps({"classes": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in c.items()}
                                      for c in ct]}), flush=True)
        batch.destroy()
        imgs.free()
        return

    for _ in range(a.warmup):
        batch.launch()
    s.synchronize()
    imgs.dst.fill(0)   # the output checked below is written by the timed launches alone
    t0 = time.perf_counter()
    for _ in range(a.steps):
        batch.launch()
    s.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    value = texels / dt / 1e6
    bit_exact = verify_batch_output(imgs, a.workload)
    log(f"batch output after the timed launches equals the reference digest: {bit_exact}")
    if bit_exact is False:
        print(json.dumps({"error": "the timed batch output differs from the reference digest",
                          "workload": a.workload}), flush=True)
        sys.exit(3)

    # per-class kernel times (HIP events on the launch stream), repeated
    classes = None
    for _ in range(max(1, min(a.steps, 3))):
        ct = batch.class_times()
        if classes is None:
            classes = ct
        else:
            for c, n in zip(classes, ct):
                c["ms"] = min(c["ms"], n["ms"])
    names = s.COLORDIST_NAMES
    for c in classes:
        class_roofline(s, c, imgs, len(classes))
    dom = max(classes, key=lambda c: c["ms"])
    kname = dom["kernel_name"]
    tr = measured_traffic().get(kname) if a.workload == "batch" else None
    hbm = {"achieved": round(dom["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom["hbm_frac"], 4),
           "algorithmic_bytes": dom["bytes"]}
    if dom["valu_ops"]:
        # all-16 and random-candidate classes: bound by VALU issue
        roofline = {"bound": "valu", "achieved": round(dom["valu_tops"], 3), "peak": VALU_PEAK_TOPS,
                    "unit": "TOP/s (int32 lane-ops)", "frac": round(dom["valu_frac"], 4),
                    "algorithmic_ops": dom["valu_ops"], "ops_basis": OPS_BASIS, "hbm": hbm}
    else:
        roofline = dict(bound="hbm", **hbm)
    roofline.update(traffic=sum(tr) if tr else None, kernel=kname, kernel_ms=round(dom["ms"], 4))
    if tr:
        roofline["traffic_source"] = "profiles/traffic_per_launch.txt (rocprofv3 FETCH_SIZE x2, WRITE_SIZE x1)"
    form_names = {0: "lane", 1: "lane_2waves", 2: "row16"}
    class_summary = [dict(kernel=c["kernel_name"], form=form_names[c["form"]], cd=names[c["cd"]], dxt=c["dxt"],
                          refine=c["refine"],
                          jobs=c["njobs"], texels=c["texels"], ms=round(c["ms"], 4), gbs=round(c["gbs"], 1),
                          hbm_frac=round(c["hbm_frac"], 4),
                          valu_frac=round(c["valu_frac"], 4) if c["valu_ops"] else None) for c in classes]

    plain = None
    if not a.no_plain:
        s.encode_jobs_plain(imgs)
        s.synchronize()
        k2 = max(1, min(a.steps, 3))
        t0 = time.perf_counter()
        for _ in range(k2):
            s.encode_jobs_plain(imgs)
        s.synchronize()
        pdt = (time.perf_counter() - t0) / k2
        plain = dict(ms_per_step=round(pdt * 1e3, 3), value=round(texels / pdt / 1e6, 1),
                     launches=len(items), speedup_batched=round(pdt / dt, 2))

    quick = None
    if not a.no_quick and a.workload == "batch":
        quick = quick_leg(s, a.quick_size)

    side = None
    if not a.no_side and a.workload == "batch":
        side = side_leg(s)

    calls = None
    if not a.no_calls and a.workload == "batch":
        calls = calls_leg(s, not a.no_cpu)

    serve = None
    if not a.no_serve and a.workload == "batch":
        serve = serve_leg(s, (2000, 10000, 50000, 60000), a.serve_duration)

    configs = None
    if not a.no_configs and a.workload == "batch":
        configs = image_configs_leg(s, not a.no_cpu)

    cpu = None
    quality = None
    if not a.no_cpu:
        cpu = cpu_baseline(items, imgs=imgs)
        if cpu:
            cpu["value"] = round(cpu["value"], 2)
        if a.workload == "batch":
            quality = quality_leg(s)

    out = {
        "metric": "Mtexels/s bit-exact, batched persistent vs per-texture plain; p99 serve latency",
        "value": round(value, 1), "unit": "Mtexels/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": {"batch": "batch: 2000 RGBA8 textures 4..2048 per side + full mip chains (16467 images), "
                                "DXT1/3/5 30/30/40%, 8 distance modes, quick selection, REFINE_ALWAYS, DITHER_NONE",
                                "servepool": "servepool: the serving pool as one batch, 256 RGBA8 textures of the "
                                "serving mix (70% quick, 30% nrandom 64 + REFINE_LOOP) + full mip chains"
                                }.get(a.workload, a.workload),
                   "images": len(items), "texels": texels, "schedule": {"wg_per_cu": sched[0], "streams": sched[1]}},
        "bit_exact": bit_exact,
        "plain_per_texture": plain,
        "device_mips": mip_leg,
        "roofline": roofline,
        "classes": class_summary,
        "quick_mode": quick,
        "side_kernels": side,
        "calls": calls,
        "cpu_baseline": cpu,
        "image_configs": configs,
        "serve": serve,
        "quality": quality,
    }
    print(json.dumps(out), flush=True)
    batch.destroy()
    imgs.free()


if __name__ == "__main__":
    main()
